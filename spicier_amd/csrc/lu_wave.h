// lu_wave.h — partial-pivot LU of one n x n system (n <= N <= 64) held by one
// 64-wide wavefront, row `lane` in lane `lane`'s registers.
//
// Layout: a[j] of lane r is A[r][j]; j is a compile-time index so the row
// stays in VGPRs (N doubles = 2N VGPRs).  Rows are never moved: each lane
// keeps `pos`, its row's current position in the partially pivoted matrix,
// and a swap of positions k <-> p reproduces the physical row swap of the
// reference's LU.  The pivot of step k is the largest |A[.][k]| over rows at
// positions >= k, smallest position on ties — nalgebra's icamax rule (first
// maximum), which CpuBatchedSolver inherits (spicier-batched-sweep
// src/solver.rs:370-380); the comparison runs on the top 14 mantissa bits,
// so candidates that close count as tied (the pivot then differs from the
// reference's only between values equal to 6e-5 relative).  Multipliers are formed as a * (1/pivot) and
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

template <typename T>
__device__ __forceinline__ T tabs(T v) { return v < T(0) ? -v : v; }

// Pivot-row broadcast: kBcastReadlane moves each element through an SGPR
// (v_readlane, VALU), kBcastLds has the pivot lane store its row to LDS and
// every lane read it back with same-address loads (LDS pipe).  The first is
// VALU-bound, the second LDS-bound; which wins depends on n and occupancy.
enum { kBcastReadlane = 0, kBcastLds = 1 };

// Doubles of the LDS buffer a RowLU<N, *> needs behind `row`: N for the
// pivot-row broadcast, and at least kWave ints for the position-map scatter.
__host__ __device__ constexpr int row_buffer_doubles(int N) { return N < 32 ? 32 : N; }

template <int N, typename T, int Bcast = kBcastReadlane>
struct RowLU {
  T a[N];
  int pos;     // position of this lane's row after pivoting
  int orig;    // lane whose row sits at position `lane`
  T dinv;      // 1 / U[pos][pos] of this lane's row
  T* row;      // N elements of LDS: broadcast buffer (and scatter space)

  // Factor in place.  Returns false if a pivot column is zero (or NaN): the
  // reference's LU then reports the system singular.  Every step runs over
  // all N columns with no per-column guard, so the update is straight-line
  // code; rows and columns past n are identity padding, whose multipliers
  // are zero.
  __device__ __forceinline__ bool factor(int lane) {
    pos = lane;
    dinv = T(0);
    bool ok = true;
    // One basic block for all N steps (no early exit, no data-dependent
    // branch), so the scheduler can start step k+1's pivot search while step
    // k's trailing updates are still in flight.  A zero pivot column clears
    // `ok`; its step then runs with a zero reciprocal and the result is
    // discarded by the caller.
    static_for<0, N>([&](auto K) __attribute__((always_inline)) {
      constexpr int k = K;
      bool nz;
      const int bp = __builtin_amdgcn_readfirstlane(pivot_pos(tabs(a[k]), pos >= k, pos, nz));
      ok = ok && nz;
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
        const T inv = pv != T(0) ? T(1) / pv : T(0);
        if (lane == p) dinv = inv;
        const T l = pos > k ? a[k] * inv : T(0);
        if (pos > k) a[k] = l;
        static_for<k + 1, N>([&](auto J) __attribute__((always_inline)) {
          constexpr int j = J;
          a[j] = fma(-row[j], l, a[j]);
        });
        __syncthreads();
      } else {
        const T pv = bcast(a[k], p);
        const T inv = pv != T(0) ? T(1) / pv : T(0);
        if (lane == p) dinv = inv;
        const T l = pos > k ? a[k] * inv : T(0);
        if (pos > k) a[k] = l;
        static_for<k + 1, N>([&](auto J) __attribute__((always_inline)) {
          constexpr int j = J;
          a[j] = fma(-bcast(a[j], p), l, a[j]);
        });
      }
    });
    if (ok) {  // invert the position map once, so solves need no ballot per step
      // all 64 lanes scatter: `row` must hold kWave ints (row_buffer_doubles)
      int* slot = reinterpret_cast<int*>(row);
      slot[pos] = lane;
      __syncthreads();
      orig = slot[lane];
      __syncthreads();
    }
    return ok;
  }

  // Solve with the stored factors for right-hand side b (b of lane r is
  // b[r]).  Returns x[lane] in each lane.  Forward substitution replays the
  // elimination on b; back substitution divides by the pivot (exact division
  // when `divide`, multiplication by the stored inverse otherwise).  The lane
  // of step k is read from `orig` of lane k, independent of b, so only the
  // broadcast of b itself is on the dependency chain.
  template <bool divide>
  __device__ __forceinline__ T solve(T b, int lane) const {
    static_for<0, N>([&](auto K) __attribute__((always_inline)) {
      constexpr int k = K;
      const int p = __builtin_amdgcn_readlane(orig, k);
      const T l = pos > k ? a[k] : T(0);
      b = fma(-bcast(b, p), l, b);
    });
    T x = T(0);
    static_for_down<0, N>([&](auto K) __attribute__((always_inline)) {
      constexpr int k = K;
      const int p = __builtin_amdgcn_readlane(orig, k);
      const T bk = bcast(b, p);
      const T xk = divide ? bk / bcast(a[k], p) : bk * bcast(dinv, p);
      if (lane == k) x = xk;
      b = fma(-(pos < k ? xk : T(0)), a[k], b);
    });
    return x;
  }
};

}  // namespace spx
