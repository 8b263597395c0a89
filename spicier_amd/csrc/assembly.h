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

// eval_sin (waveforms.rs:185-201): phase in degrees, damping from TD on.
__device__ __forceinline__ double eval_sin(double vo, double va, double freq, double td,
                                           double theta, double phase, double t) {
  if (t < td) return vo;
  const double t_rel = t - td;
  const double phase_rad = phase * M_PI / 180.0;
  const double damping = theta > 0.0 ? exp(-theta * t_rel) : 1.0;
  return vo + va * damping * sin(2.0 * M_PI * freq * t_rel + phase_rad);
}

// eval_pwl (waveforms.rs:204-232) over np (t, v) pairs in consecutive
// parameter slots from `first`: the first value up to t1, the last from tk
// on, else linear interpolation in the first interval [t_i, t_i+1] that
// holds t.
template <int ST>
__device__ __forceinline__ double eval_pwl(const double* pv, int first, int np, double t) {
  auto T = [&](int i) { return pv[(first + 2 * i) * ST]; };
  auto V = [&](int i) { return pv[(first + 2 * i + 1) * ST]; };
  if (t <= T(0)) return V(0);
  if (t >= T(np - 1)) return V(np - 1);
  for (int i = 0; i < np - 1; ++i) {
    const double t0 = T(i), t1 = T(i + 1);
    if (t >= t0 && t <= t1) {
      const double frac = (t - t0) / (t1 - t0);
      return V(i) + frac * (V(i + 1) - V(i));
    }
  }
  return V(np - 1);
}

// Companion values of element e at node voltages xs (unknown u -> xs[u],
// ground -> 0): conductances, currents and couplings into vals[vid..].
// sf scales independent sources (source stepping); pnjlim junction state
// lives in vst[e]; `lim` is set when a pnjlim limiter acted.  tc is the
// transient context (null for a DC sweep, whose program has no capacitor
// slots).
//
// ST is the stride of every per-instance array (xs, pv, vals, vst and the
// transient's vprev): 1 when one workgroup owns an instance, 64 when the 64
// lanes of a wave are 64 instances with their values interleaved.  ES is
// the stride between the element record's fields: 1 in elems, 64 in the
// lane-major el_lm.
template <int ST = 1, int ES = 1>
__device__ __forceinline__ void eval_element_at(const int* rec, int e, const double* xs,
                                                const double* pv, double* vals, double* vst,
                                                double sf, int limiter, bool& lim,
                                                const TranCtx* tc = nullptr) {
  auto F = [&](int k) { return rec[k * ES]; };
  const int kind = F(0), vid = F(11);
  auto X = [&](int u) { return u >= 0 ? xs[u * ST] : 0.0; };
  auto PV = [&](int k) { return pv[F(k) * ST]; };
  switch (kind) {
    case SPX_EL_RESISTOR:
      vals[vid * ST] = 1.0 / PV(4);
      break;
    case SPX_EL_CAPACITOR:  // open at DC; the stage's companion (companion.rs:41-130)
      if (vid >= 0) {
        double geq = 0.0, ieq = 0.0;
        if (tc && tc->mode == 1) {
          const double C = PV(4);
          if (tc->comp == 1) {
            geq = 2.0 * C / tc->h;
            ieq = geq * tc->vprev[e * ST] + tc->iprev[e * ST];
          } else if (tc->comp == 2) {
            geq = C / tc->bden;
            ieq = geq * (tc->a1 * tc->vprev[e * ST] + tc->a2 * tc->vpp[e * ST]);
          } else {
            geq = C / tc->h;
            ieq = geq * tc->vprev[e * ST];
          }
        }
        vals[vid * ST] = geq;
        vals[(vid + 1) * ST] = ieq;
      }
      break;
    case SPX_EL_INDUCTOR:  // branch row c (v0 - v1) + d i = r
      if (tc && tc->mode == 1) {
        const double L = PV(4);
        double geq, ieq;
        if (tc->comp == 1) {
          geq = tc->h / (2.0 * L);
          ieq = tc->iprev[e * ST] + geq * tc->vprev[e * ST];
        } else if (tc->comp == 2) {
          geq = tc->bden / L;
          ieq = tc->a1 * tc->iprev[e * ST] + tc->a2 * tc->vpp[e * ST];
        } else {
          geq = tc->h / L;
          ieq = tc->iprev[e * ST];
        }
        vals[vid * ST] = -geq;
        vals[(vid + 1) * ST] = 1.0;
        vals[(vid + 2) * ST] = ieq;
      } else {  // 0 V source
        vals[vid * ST] = 1.0;
        vals[(vid + 1) * ST] = 0.0;
        vals[(vid + 2) * ST] = 0.0;
      }
      break;
    case SPX_EL_VPWL:  // field 3: number of points, field 4: the slot of t1
      vals[vid * ST] = ((tc && tc->mode == 1) ? eval_pwl<ST>(pv, F(4), F(3), tc->t)
                                              : pv[(F(4) + 1) * ST]) *
                       sf;
      break;
    case SPX_EL_VSIN:
      vals[vid * ST] =
          ((tc && tc->mode == 1) ? eval_sin(PV(4), PV(5), PV(6), PV(7), PV(8), PV(9), tc->t)
                                 : PV(4)) *
          sf;
      break;
    case SPX_EL_VPULSE:
      vals[vid * ST] =
          ((tc && tc->mode == 1) ? eval_pulse(PV(4), PV(5), PV(6), PV(7), PV(8), PV(9), PV(10),
                                              tc->t)
                                 : PV(4)) *
          sf;
      break;
    case SPX_EL_VSOURCE:
    case SPX_EL_ISOURCE:
      vals[vid * ST] = PV(4) * sf;
      break;
    case SPX_EL_VCVS:  // gains: dependent sources do not take source stepping
    case SPX_EL_VCCS:
    case SPX_EL_CCCS:
    case SPX_EL_CCVS:
      vals[vid * ST] = PV(4);
      break;
    case SPX_EL_DIODE: {
      double vd = X(F(1)) - X(F(2));
      double gd, ieq;
      diode_companion(PV(4), PV(5), vd, limiter, vst[e * ST], lim, gd, ieq);
      vals[vid * ST] = gd;
      vals[(vid + 1) * ST] = ieq;
      break;
    }
    case SPX_EL_NMOS:
    case SPX_EL_PMOS: {
      double vd = X(F(1)), vg = X(F(2)), vs = X(F(3));
      double gds, gm, ieq;
      mosfet_companion(kind == SPX_EL_PMOS, PV(4), PV(5), PV(6), PV(7), PV(8), vg - vs, vd - vs,
                       gds, gm, ieq);
      vals[vid * ST] = gds;
      vals[(vid + 1) * ST] = gm;
      vals[(vid + 2) * ST] = ieq;
      break;
    }
    default:
      break;
  }
}

template <int ST = 1>
__device__ __forceinline__ void eval_element(const DevProgram& g, int e, const double* xs,
                                             const double* pv, double* vals, double* vst,
                                             double sf, int limiter, bool& lim,
                                             const TranCtx* tc = nullptr) {
  eval_element_at<ST, 1>(g.elems + e * kElemInts, e, xs, pv, vals, vst, sf, limiter, lim, tc);
}

// The pattern and rhs sums of one 64-lane wave from the lane-major gather
// program (program.h): the same contributions in the same order as
// gather_sum, with the op loads independent of each other.
__device__ __forceinline__ void gather_lanes(const DevProgram& g, int lane, double* J,
                                             double* rhs, const double* vals) {
  double acc = 0.0;
  const int* ops = g.gath + lane;
#pragma unroll 4
  for (int k = 0; k < g.gath_k; ++k) {
    const int op = ops[k * 64];
    if (op != -1) {
      const double v = vals[op & 0x3fff];
      acc = (op & 0x4000) ? acc - v : acc + v;
      if (op & 0x8000) {
        const unsigned d = (unsigned)op >> 16;
        if (d & 0x8000u)
          rhs[d & 0x7fffu] = acc;
        else
          J[d] = acc;
        acc = 0.0;
      }
    }
  }
}

// Sum of entry q's contributions (value slot << 1 | negate), in push order.
template <int ST = 1>
__device__ __forceinline__ double gather_sum(const int* off, const int* contrib, int q,
                                             const double* vals) {
  double sum = 0.0;
  const int c1 = off[q + 1];
  for (int c = off[q]; c < c1; ++c) {
    const int t = contrib[c];
    const double v = vals[(t >> 1) * ST];
    sum = (t & 1) ? sum - v : sum + v;
  }
  return sum;
}

}  // namespace spx
