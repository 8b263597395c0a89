// assembly.h — device evaluation and pattern assembly shared by the Newton
// kernels: one element's companion model into its value slots, and one
// pattern entry's (or rhs row's) sum of contributions in the reference's
// triplet push order (mna.rs:62-71, summed as to_dense_matrix does).
#pragma once

#include "device_common.h"
#include "program.h"

namespace spx {

// PULSE source value at time t (waveforms.rs:160-180).
__device__ __forceinline__ double eval_pulse(double v1, double v2, double td, double tr,
                                             double tf, double pw, double per, double t) {
  if (t < td) return v1;
  const double tr_ = per > 0.0 ? fmod(t - td, per) : t - td;
  if (tr_ < tr) return v1 + (v2 - v1) * tr_ / tr;
  if (tr_ < tr + pw) return v2;
  if (tr_ < tr + pw + tf) return v2 - (v2 - v1) * (tr_ - tr - pw) / tf;
  return v1;
}

// Companion values of element e at node voltages xs (unknown u -> xs[u],
// ground -> 0): conductances, currents and couplings into vals[vid..].
// sf scales independent sources (source stepping); pnjlim junction state
// lives in vst[e]; `lim` is set when a pnjlim limiter acted.  tc is the
// transient context (null for a DC sweep, whose program has no capacitor
// slots).
__device__ __forceinline__ void eval_element(const DevProgram& g, int e, const double* xs,
                                             const double* pv, double* vals, double* vst,
                                             double sf, int limiter, bool& lim,
                                             const TranCtx* tc = nullptr) {
  const int* el = g.elems + e * kElemInts;
  const int kind = el[0], vid = el[11];
  auto X = [&](int u) { return u >= 0 ? xs[u] : 0.0; };
  switch (kind) {
    case SPX_EL_RESISTOR:
      vals[vid] = 1.0 / pv[el[4]];
      break;
    case SPX_EL_CAPACITOR:  // companion.rs:41-48
      if (vid >= 0) {
        const double geq = (tc && tc->mode == 1) ? pv[el[4]] / tc->h : 0.0;
        vals[vid] = geq;
        vals[vid + 1] = (tc && tc->mode == 1) ? geq * tc->vprev[e] : 0.0;
      }
      break;
    case SPX_EL_VPULSE:
      vals[vid] = ((tc && tc->mode == 1)
                       ? eval_pulse(pv[el[4]], pv[el[5]], pv[el[6]], pv[el[7]], pv[el[8]],
                                    pv[el[9]], pv[el[10]], tc->t)
                       : pv[el[4]]) *
                  sf;
      break;
    case SPX_EL_VSOURCE:
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
