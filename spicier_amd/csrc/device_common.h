// device_common.h — device-side building blocks of the sweep engine:
// hash RNG, the LCG generator with jump-ahead, device models, and wave-level
// reductions for 64-wide CDNA4 wavefronts.
//
// Arithmetic order follows the reference so that, with -ffp-contract=off,
// device models round as the reference's CPU code does; the LU kernels use
// explicit fma() where fusing is wanted.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/spicier_hip.h"

namespace spx {

constexpr int kWave = 64;
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt2 = 1.41421356237309504880;

// ------------------------------------------------------------------ rng.rs
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {  // rng.rs:39-44
  x = x + 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

__device__ __forceinline__ double hash_uniform(uint64_t seed, uint32_t s, uint32_t p) {
  // rng.rs:51-56 combine_indices + rng.rs:80-85 top 53 bits
  uint64_t h = splitmix64(seed ^ ((uint64_t)s * 0x517cc1b727220a95ULL) ^
                          ((uint64_t)p * 0x5851f42d4c957f2dULL));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double hash_gaussian(uint64_t seed, uint32_t s, uint32_t p) {
  // rng.rs:120-131: Box-Muller cos branch on params 2p, 2p+1
  double u1 = hash_uniform(seed, s, p * 2u);
  double u2 = hash_uniform(seed, s, p * 2u + 1u);
  u1 = fmax(u1, 1e-10);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

// MonteCarloGenerator's LCG (spicier-solver/src/sweep.rs:173-178) advanced by
// k steps in O(log k): the affine map s -> a*s + c composed by squaring.
__device__ __forceinline__ uint64_t lcg_skip(uint64_t s, uint64_t k) {
  uint64_t am = 1, ac = 0, cm = 6364136223846793005ULL, cc = 1;
  while (k) {
    if (k & 1) { am *= cm; ac = ac * cm + cc; }
    cc = (cm + 1) * cc;
    cm *= cm;
    k >>= 1;
  }
  return am * s + ac;
}

struct Lcg {
  uint64_t s;
  __device__ double next_random() {
    s = s * 6364136223846793005ULL + 1ULL;
    return (double)(s >> 33) / 2147483648.0;
  }
  __device__ double next_normal() {  // sweep.rs:181-185
    double u1 = fmax(next_random(), 1e-10);
    double u2 = next_random();
    return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
  }
};

__device__ __forceinline__ double rclamp(double v, double lo, double hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

// ------------------------------------------------------------------ devices
// diode.rs:41-45 at 300.15 K
__device__ __forceinline__ double thermal_voltage() {
  return 1.380649e-23 * 300.15 / 1.602176634e-19;
}

// diode.rs:139-148 (IS fixed at 1e-14 inside vcrit, as the reference)
__device__ __forceinline__ double diode_limit_ref(double vd, double nvt) {
  double vcrit = nvt * log(nvt / (kSqrt2 * 1e-14));
  if (vd > vcrit) {
    double arg = (vd - vcrit) / nvt;
    return vcrit + nvt * (1.0 + log1p(arg));
  }
  return vd;
}

// SPICE3 DEVpnjlim
__device__ __forceinline__ double pnjlim(double vnew, double vold, double vt, double vcrit,
                                         bool& limited) {
  if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {
    if (vold > 0) {
      double arg = 1 + (vnew - vold) / vt;
      vnew = arg > 0 ? vold + vt * log(arg) : vcrit;
    } else {
      vnew = vt * log(vnew / vt);
    }
    limited = true;
  }
  return vnew;
}

// Linearised diode (diode.rs:91-131): conductance gd and companion current
// ieq at junction voltage vd; vstate is the pnjlim memory.
__device__ __forceinline__ void diode_companion(double is, double n, double vd, int limiter,
                                                double& vstate, bool& limited, double& gd,
                                                double& ieq) {
  double nvt = n * thermal_voltage();
  double vl, e;
  if (limiter == SPX_LIMIT_PNJLIM) {
    double vcrit = nvt * log(nvt / (kSqrt2 * is));
    vl = pnjlim(vd, vstate, nvt, vcrit, limited);
    vstate = vl;
    e = exp(vl / nvt);
  } else if (limiter == SPX_LIMIT_REFERENCE) {
    vl = diode_limit_ref(vd, nvt);
    e = exp(diode_limit_ref(vl, nvt) / nvt);  // evaluate() limits again (diode.rs:96)
  } else {
    vl = vd;
    e = exp(vl / nvt);
  }
  double id = is * (e - 1.0);
  gd = fmax(is * e / nvt, 1e-12);
  ieq = id - gd * vl;
}

// level1.rs:154-191 + 243: ids, gds, gm and the companion current.
__device__ __forceinline__ void mosfet_companion(bool pmos, double vto, double kp, double lambda,
                                                 double w, double l, double vgs0, double vds0,
                                                 double& gds, double& gm, double& ieq) {
  double vgs = pmos ? -vgs0 : vgs0, vds = pmos ? -vds0 : vds0;
  double vth = fabs(vto);
  double beta = kp * 1.0 * w / l;
  double ids, gd, g;
  if (vgs < vth) {
    ids = 0.0; gd = 1e-12; g = 0.0;
  } else if (vds < vgs - vth) {
    double vov = vgs - vth;
    double i = beta * (vov * vds - 0.5 * vds * vds) * (1.0 + lambda * vds);
    gd = beta * (vov - vds) * (1.0 + lambda * vds) + beta * (vov * vds - 0.5 * vds * vds) * lambda;
    g = beta * vds * (1.0 + lambda * vds);
    ids = pmos ? -i : i;
    gd = fmax(gd, 1e-12);
  } else {
    double vov = vgs - vth;
    double i = 0.5 * beta * vov * vov * (1.0 + lambda * vds);
    gd = 0.5 * beta * vov * vov * lambda;
    g = beta * vov * (1.0 + lambda * vds);
    ids = pmos ? -i : i;
    gd = fmax(gd, 1e-12);
  }
  gds = gd;
  gm = g;
  ieq = ids - gds * vds0 - gm * vgs0;
}

// ------------------------------------------------------------- wave helpers
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = __builtin_amdgcn_readlane(p.x, lane);
  p.y = __builtin_amdgcn_readlane(p.y, lane);
  return *reinterpret_cast<double*>(&p);
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// One DPP max step on a 64-bit key (two 32-bit moves; lanes whose source is
// outside the pattern read 0, the identity of an unsigned max).
template <int Ctrl, int RowMask>
__device__ __forceinline__ uint64_t dpp_max_u64(uint64_t v) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, Ctrl, RowMask, 0xf, false);
  uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), Ctrl, RowMask, 0xf,
                                                      false);
  uint64_t o = ((uint64_t)hi << 32) | lo;
  return o > v ? o : v;
}

template <int Ctrl, int RowMask>
__device__ __forceinline__ uint32_t dpp_max_u32(uint32_t v) {
  uint32_t o = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, RowMask, 0xf, false);
  return o > v ? o : v;
}

// Wave-wide unsigned max with DPP (quad_perm [1,0,3,2], [2,3,0,1], row_shr:4,
// row_shr:8, row_bcast:15, row_bcast:31 — the GFX9 reduction ladder), the
// result read from lane 63 so it is wave-uniform.  No LDS round trips.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  v = dpp_max_u64<0xb1, 0xf>(v);
  v = dpp_max_u64<0x4e, 0xf>(v);
  v = dpp_max_u64<0x114, 0xf>(v);
  v = dpp_max_u64<0x118, 0xf>(v);
  v = dpp_max_u64<0x142, 0xa>(v);
  v = dpp_max_u64<0x143, 0xc>(v);
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = dpp_max_u32<0xb1, 0xf>(v);
  v = dpp_max_u32<0x4e, 0xf>(v);
  v = dpp_max_u32<0x114, 0xf>(v);
  v = dpp_max_u32<0x118, 0xf>(v);
  v = dpp_max_u32<0x142, 0xa>(v);
  v = dpp_max_u32<0x143, 0xc>(v);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Pivot choice for the register LU: one 32-bit DPP max over
// key = (high word of |a| with its 6 low bits cleared) | (63 - pos).  The
// largest |a| wins; candidates equal in sign, exponent and the top 14
// mantissa bits — exact ties included — go to the smallest position, which is
// nalgebra's first-maximum rule for exact ties.  Returns the winning position
// (pos < 64) and whether the column held a non-zero value.
__device__ __forceinline__ int pivot_pos(double absval, bool eligible, int pos, bool& nonzero) {
  const uint32_t hi = (uint32_t)((uint64_t)__double_as_longlong(absval) >> 32);
  const uint32_t key = eligible ? ((hi & ~63u) | (uint32_t)(63 - pos)) : 0u;
  const uint32_t mx = wave_max_u32(key);
  nonzero = (mx >> 6) != 0;
  return 63 - (int)(mx & 63u);
}

// 1 if this lane's candidate a could change the pivot pv that pivot_pos
// chose under the exact rule: eligible, equal to |pv| in the compared bits
// (sign cleared, the high word's 6 low bits dropped) but not in value.
// Integer compares only, so that the caller accumulates the result in a VGPR
// across the elimination without holding lane masks.
__device__ __forceinline__ uint32_t pivot_inexact(double a, double pv, bool eligible) {
  const uint64_t ba = (uint64_t)__double_as_longlong(a), bp = (uint64_t)__double_as_longlong(pv);
  const uint32_t hi = (uint32_t)(ba >> 32) & 0x7fffffffu, hp = (uint32_t)(bp >> 32) & 0x7fffffffu;
  const uint32_t lo = (uint32_t)ba, lp = (uint32_t)bp;
  return (eligible && (hi ^ hp) < 64u && (hi != hp || lo != lp)) ? 1u : 0u;
}

// The exact rule: the strictly largest |a| over every bit, then the
// smallest position.  After the truncated pass, lanes tied on its key
// compare the remaining 38 bits of |a| (the high word's 6 low bits with the
// low word's top 26, then the low word's last 6 bits with 63 - pos) in two
// more DPP max passes, which run only on such ties (a wave-uniform branch).
// A denormal pivot counts as non-zero.
__device__ __forceinline__ int pivot_pos_exact(double absval, bool eligible, int pos,
                                               bool& nonzero) {
  const uint64_t bits = (uint64_t)__double_as_longlong(absval);
  const uint32_t hi = (uint32_t)(bits >> 32), lo = (uint32_t)bits;
  const uint32_t key = eligible ? ((hi & ~63u) | (uint32_t)(63 - pos)) : 0u;
  const uint32_t mx = wave_max_u32(key);
  const bool tied = eligible && (key >> 6) == (mx >> 6);
  if (__popcll(__ballot(tied)) > 1) {
    const uint32_t ka = tied ? (((hi & 63u) << 26) | (lo >> 6)) : 0u;
    const uint32_t ma = wave_max_u32(ka);
    const bool tied2 = tied && ka == ma;
    const uint32_t kb = tied2 ? (((lo & 63u) << 6) | (uint32_t)(63 - pos)) : 0u;
    const uint32_t mb = wave_max_u32(kb);
    nonzero = (mx >> 6) != 0 || ma != 0 || (mb >> 6) != 0;
    return 63 - (int)(mb & 63u);
  }
  nonzero = (mx >> 6) != 0;
  return 63 - (int)(mx & 63u);
}

// FP32 keeps the truncated comparison (17 mantissa bits): FP32 factors serve
// only as the preconditioner of FP64 iterative refinement, whose stopping
// test, not the pivot order, decides the result.
__device__ __forceinline__ int pivot_pos(float absval, bool eligible, int pos, bool& nonzero) {
  const uint32_t bits = __float_as_uint(absval);
  const uint32_t key = eligible ? ((bits & ~63u) | (uint32_t)(63 - pos)) : 0u;
  const uint32_t mx = wave_max_u32(key);
  nonzero = (mx >> 6) != 0;
  return 63 - (int)(mx & 63u);
}

// One DPP max step on doubles that are >= 0 (or NaN): lanes whose source is
// outside the pattern read +0.0, which leaves the max unchanged.
template <int Ctrl, int RowMask>
__device__ __forceinline__ double dpp_max_nonneg(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, Ctrl, RowMask,
                                                            0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), Ctrl,
                                                            RowMask, 0xf, false);
  return fmax(v, __longlong_as_double((long long)(((uint64_t)hi << 32) | lo)));
}

// Wave-wide max of values >= 0 (NaN lanes are skipped, as fmax skips them),
// on the DPP ladder of wave_max_u32: no LDS round trips.
__device__ __forceinline__ double wave_max(double v) {
  v = dpp_max_nonneg<0xb1, 0xf>(v);
  v = dpp_max_nonneg<0x4e, 0xf>(v);
  v = dpp_max_nonneg<0x114, 0xf>(v);
  v = dpp_max_nonneg<0x118, 0xf>(v);
  v = dpp_max_nonneg<0x142, 0xa>(v);
  v = dpp_max_nonneg<0x143, 0xc>(v);
  return readlane_f64(v, 63);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

}  // namespace spx
