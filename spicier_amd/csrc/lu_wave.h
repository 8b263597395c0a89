// lu_wave.h — partial-pivot LU of one n x n system (n <= N <= 64) held by one
// 64-wide wavefront, row `lane` in lane `lane`'s registers.
//
// Layout: a[j] of lane r is A[r][j]; j is a compile-time index so the row
// stays in VGPRs (N doubles = 2N VGPRs).  Rows do not move during the
// elimination: each lane keeps `pos`, its row's current position in the
// partially pivoted matrix, and a swap of positions k <-> p reproduces the
// physical row swap of the reference's LU.  After the last step every row is
// pushed to the lane of its final position (one cross-lane permute per
// register), so lane r holds row r of P A = L U and the solves address the
// row of step k as lane k, a compile-time constant.  The pivot of step k is
// the largest |A[.][k]| over rows at positions >= k, smallest position on
// exact ties — nalgebra's icamax rule (first maximum), which CpuBatchedSolver
// inherits (spicier-batched-sweep src/solver.rs:370-380).  The fast pass
// compares 14 mantissa bits (FP64) or 17 (FP32) with the position; in FP64
// it also notes whether a candidate equal to the pivot in those bits but not
// in value existed (off the steps' critical path), and `inexact` then tells
// the caller to reload the rows and factor again with Exact = true, which
// compares every bit — so FP64 pivots follow the rule exactly.  FP32
// factors keep the fast pass: they only precondition FP64 refinement.
// Multipliers are formed as a * (1/pivot) and
// updates as fma(-u, l, a).  The caller's workgroup is exactly one wave, so
// __syncthreads() orders LDS traffic without a real barrier.
//
// Lanes >= n carry identity rows (a[j] = [j == lane]) so that padding never
// wins a pivot search and solves to zero.

#pragma once

#include <type_traits>

#include "device_common.h"

namespace spx {

// Compile-time loop: f(integral_constant<int, I>) for I in [Begin, End).  The
// index is a constant expression in the body, so register arrays indexed by
// it are split into scalars before any loop optimisation runs.
template <int Begin, int End, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (Begin < End) {
    f(std::integral_constant<int, Begin>{});
    static_for<Begin + 1, End>(f);
  }
}

// Same, counting down from Hi-1 to Lo.
template <int Lo, int Hi, typename F>
__device__ __forceinline__ void static_for_down(F&& f) {
  if constexpr (Lo < Hi) {
    f(std::integral_constant<int, Hi - 1>{});
    static_for_down<Lo, Hi - 1>(f);
  }
}

__device__ __forceinline__ double bcast(double v, int lane) { return readlane_f64(v, lane); }
__device__ __forceinline__ float bcast(float v, int lane) { return readlane_f32(v, lane); }

// Cross-lane moves through the LDS crossbar (no LDS storage): push_to sends
// this lane's v to lane `dst` (dst must be a permutation of the lanes),
// pull_from returns v of lane `src`.
__device__ __forceinline__ int push_to(int v, int dst) { return __builtin_amdgcn_ds_permute(dst << 2, v); }
__device__ __forceinline__ float push_to(float v, int dst) {
  return __int_as_float(push_to(__float_as_int(v), dst));
}
__device__ __forceinline__ double push_to(double v, int dst) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = push_to(p.x, dst);
  p.y = push_to(p.y, dst);
  return *reinterpret_cast<double*>(&p);
}
__device__ __forceinline__ int pull_from(int v, int src) {
  return __builtin_amdgcn_ds_bpermute(src << 2, v);
}
__device__ __forceinline__ float pull_from(float v, int src) {
  return __int_as_float(pull_from(__float_as_int(v), src));
}
__device__ __forceinline__ double pull_from(double v, int src) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = pull_from(p.x, src);
  p.y = pull_from(p.y, src);
  return *reinterpret_cast<double*>(&p);
}

// |v| with the sign bit cleared.  The pivot search ranks the bit patterns of
// |v|, so a -0.0 (an elimination update fma(-u, l, +0.0) whose exact
// product underflows rounds to -0.0) must rank as zero, not above every
// positive value.
template <typename T>
__device__ __forceinline__ T tabs(T v) { return fabs(v); }

// Pivot-row broadcast: kBcastReadlane moves each element through an SGPR
// (v_readlane, VALU), kBcastLds has the pivot lane store its row to LDS and
// every lane read it back with same-address loads (LDS pipe).  The first is
// VALU-bound, the second LDS-bound; which wins depends on n and occupancy.
enum { kBcastReadlane = 0, kBcastLds = 1 };

// Doubles of the LDS buffer a RowLU<N, *> needs behind `row`: N for the
// pivot-row broadcast of the LDS variant (the callers reserve at least 32).
__host__ __device__ constexpr int row_buffer_doubles(int N) { return N < 32 ? 32 : N; }

template <int N, typename T, int Bcast = kBcastReadlane>
struct RowLU {
  T a[N];
  int pos;     // position of this lane's row (== lane once factored)
  int orig;    // row of the input whose factor row lane `lane` holds
  T piv;       // U[lane][lane] once factored
  T dinv;      // 1 / U[lane][lane]
  T* row;      // N elements of LDS: broadcast buffer of the LDS variant
  int ipiv;    // factor<true>: lane k holds the position swapped with k at step k
  bool inexact;  // FP64 fast pass: a near-tie needs factor<*, true> on reloaded rows

  // Factor in place.  Returns false if a pivot column is zero (or NaN): the
  // reference's LU then reports the system singular.  Every step runs over
  // all N columns with no per-column guard, so the update is straight-line
  // code; rows and columns past n are identity padding, whose multipliers
  // are zero.
  template <bool Record = false, bool Exact = false>
  __device__ __forceinline__ bool factor(int lane) {
    constexpr bool kF64 = std::is_same<T, double>::value;
    pos = lane;
    if constexpr (Record) ipiv = lane;
    dinv = T(0);
    piv = T(0);
    bool ok = true;
    uint32_t nearby = 0;  // FP64 fast pass: a near-tie the exact rule might decide otherwise

    // One basic block for all N steps (no early exit, no data-dependent
    // branch), so the scheduler can start step k+1's pivot search while step
    // k's trailing updates are still in flight.  A zero pivot column clears
    // `ok`; its step then runs with a zero reciprocal and the result is
    // discarded by the caller.
    static_for<0, N>([&](auto K) __attribute__((always_inline)) {
      constexpr int k = K;
      bool nz;
      int bp0;
      if constexpr (kF64 && Exact)
        bp0 = pivot_pos_exact(tabs(a[k]), pos >= k, pos, nz);
      else
        bp0 = pivot_pos(tabs(a[k]), pos >= k, pos, nz);
      const int bp = __builtin_amdgcn_readfirstlane(bp0);
      ok = ok && nz;
      if constexpr (Record) {
        if (lane == k) ipiv = bp;
      }
      const int p = __builtin_amdgcn_readfirstlane(
          (int)__ffsll((unsigned long long)__ballot(pos == bp)) - 1);
      if (pos == k) pos = bp;  // row at position k moves to the pivot's position
      if (lane == p) pos = k;
      if constexpr (Bcast == kBcastLds) {
        if (lane == p) {
          static_for<k, N>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = J;
            row[j] = a[j];
          });
        }
        __syncthreads();
        const T pv = row[k];
        if constexpr (kF64 && !Exact) nearby |= pivot_inexact(a[k], pv, pos >= k);
        const T inv = pv != T(0) ? T(1) / pv : T(0);
        if (lane == p) {
          dinv = inv;
          piv = pv;
        }
        const T l = pos > k ? a[k] * inv : T(0);
        if (pos > k) a[k] = l;
        static_for<k + 1, N>([&](auto J) __attribute__((always_inline)) {
          constexpr int j = J;
          a[j] = fma(-row[j], l, a[j]);
        });
        __syncthreads();
      } else {
        const T pv = bcast(a[k], p);
        if constexpr (kF64 && !Exact) nearby |= pivot_inexact(a[k], pv, pos >= k);
        // FP32 factors only precondition FP64 refinement: the hardware
        // reciprocal (1 ulp) serves; FP64 keeps the correctly rounded 1/pivot
        T inv;
        if constexpr (kF64)
          inv = pv != T(0) ? T(1) / pv : T(0);
        else
          inv = pv != T(0) ? __builtin_amdgcn_rcpf(pv) : T(0);
        if (lane == p) {
          dinv = inv;
          piv = pv;
        }
        const T l = pos > k ? a[k] * inv : T(0);
        if (pos > k) a[k] = l;
        static_for<k + 1, N>([&](auto J) __attribute__((always_inline)) {
          constexpr int j = J;
          a[j] = fma(-bcast(a[j], p), l, a[j]);
        });
      }
    });
    inexact = kF64 && !Exact && __any(nearby != 0u);
    if (inexact) return false;  // the caller reloads the rows and runs factor<*, true>
    if (ok) {  // pos is a permutation of the 64 lanes: push every row home
      orig = push_to(lane, pos);
      static_for<0, N>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = J;
        a[j] = push_to(a[j], pos);
      });
      dinv = push_to(dinv, pos);
      piv = push_to(piv, pos);
      pos = lane;
    }
    return ok;
  }

  // Solve with the stored factors for right-hand side b (b of lane r is
  // b[r]).  Returns x[lane] in each lane.  b is first pulled into pivoted
  // order; forward substitution then runs the elimination's updates on it
  // (lanes below the diagonal, L[r][k] = a[k]); back substitution divides by
  // the pivot (exact division when `divide`, multiplication by the stored
  // inverse otherwise) and updates the rows above it.  Step k reads lane k,
  // so the only cross-lane read on the dependency chain is that broadcast.
  // x of lane k is the value step k broadcast: the last update of lane k's b
  // happens before step k, so one vector divide (or multiply) at the end
  // reproduces it bit for bit.
  template <bool divide>
  __device__ __forceinline__ T solve(T b, int lane) const {
    b = pull_from(b, orig);
    // a fresh value per call, so the per-step lane tests are not hoisted out
    // of a caller's loop as 2N live lane masks
    int ln = lane;
    asm volatile("" : "+v"(ln));
    static_for<0, N - 1>([&](auto K) __attribute__((always_inline)) {
      constexpr int k = K;
      const T yk = bcast(b, k);
      if (ln > k) b = fma(-yk, a[k], b);
    });
    static_for_down<1, N>([&](auto K) __attribute__((always_inline)) {
      constexpr int k = K;
      T xk;
      if constexpr (divide) {
        xk = bcast(b, k) / bcast(piv, k);
      } else {
        xk = bcast(b * dinv, k);
      }
      if (ln < k) b = fma(-xk, a[k], b);
    });
    const T x = divide ? b / piv : b * dinv;
    return lane < N ? x : T(0);  // lanes past N hold no row (piv = 0)
  }
};

}  // namespace spx
