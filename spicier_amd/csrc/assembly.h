// assembly.h — device evaluation and pattern assembly shared by the Newton
// kernels: one element's companion model into its value slots, and one
// pattern entry's (or rhs row's) sum of contributions in the reference's
// triplet push order (mna.rs:62-71, summed as to_dense_matrix does).
#pragma once

#include "device_common.h"
#include "program.h"

namespace spx {

// Companion values of element e at node voltages xs (unknown u -> xs[u],
// ground -> 0): conductances, currents and couplings into vals[vid..].
// sf scales independent sources (source stepping); pnjlim junction state
// lives in vst[e]; `lim` is set when a pnjlim limiter acted.
__device__ __forceinline__ void eval_element(const DevProgram& g, int e, const double* xs,
                                             const double* pv, double* vals, double* vst,
                                             double sf, int limiter, bool& lim) {
  const int* el = g.elems + e * kElemInts;
  const int kind = el[0], vid = el[11];
  auto X = [&](int u) { return u >= 0 ? xs[u] : 0.0; };
  switch (kind) {
    case SPX_EL_RESISTOR:
      vals[vid] = 1.0 / pv[el[4]];
      break;
    case SPX_EL_VSOURCE:
    case SPX_EL_VPULSE:
    case SPX_EL_ISOURCE:
      vals[vid] = pv[el[4]] * sf;
      break;
    case SPX_EL_DIODE: {
      double vd = X(el[1]) - X(el[2]);
      double gd, ieq;
      diode_companion(pv[el[4]], pv[el[5]], vd, limiter, vst[e], lim, gd, ieq);
      vals[vid] = gd;
      vals[vid + 1] = ieq;
      break;
    }
    case SPX_EL_NMOS:
    case SPX_EL_PMOS: {
      double vd = X(el[1]), vg = X(el[2]), vs = X(el[3]);
      double gds, gm, ieq;
      mosfet_companion(kind == SPX_EL_PMOS, pv[el[4]], pv[el[5]], pv[el[6]], pv[el[7]],
                       pv[el[8]], vg - vs, vd - vs, gds, gm, ieq);
      vals[vid] = gds;
      vals[vid + 1] = gm;
      vals[vid + 2] = ieq;
      break;
    }
    default:
      break;
  }
}

// Sum of entry q's contributions (value slot << 1 | negate), in push order.
__device__ __forceinline__ double gather_sum(const int* off, const int* contrib, int q,
                                             const double* vals) {
  double sum = 0.0;
  const int c1 = off[q + 1];
  for (int c = off[q]; c < c1; ++c) {
    const int t = contrib[c];
    const double v = vals[t >> 1];
    sum = (t & 1) ? sum - v : sum + v;
  }
  return sum;
}

}  // namespace spx
