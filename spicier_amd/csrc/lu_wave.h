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

// Columns are grouped in chunks of kChunk for the structural skip of the
// elimination updates (one bit per chunk in RowLU::factor's masks).
constexpr int kChunk = 8;

// Chunk mask of a row held in registers: bit c set if any of columns
// [c*kChunk, (c+1)*kChunk) is non-zero.
template <int N, typename T>
__device__ __forceinline__ uint32_t chunk_mask_of(const T (&a)[N]) {
  uint32_t m = 0;
  static_for<0, N>([&](auto J) __attribute__((always_inline)) {
    constexpr int j = J;
    if (a[j] != T(0)) m |= 1u << (j / kChunk);
  });
  return m;
}

// One chunk of the FP32 elimination update a[j] = fma(-u_j, l, a[j]) for
// j in [J0, J0 + CN), u_j = a[j] of lane p, skipped as a whole when bit C of
// the pivot row's chunk mask mp is clear.  The test and the skip are inside
// one asm statement, so the compiler sees straight-line code (an update
// branch per chunk written in C++ splits the unrolled elimination into
// hundreds of blocks, and the register allocator then spills the rows).
// The CN broadcasts are issued before the CN FMAs, which covers the
// SGPR-write-to-VALU-read wait of each; chunks of fewer than three columns
// add it explicitly.  The leading s_nop covers a VALU write of mp just
// before the SALU test.
#define SPX_RL(i) "v_readlane_b32 %[t" #i "], %[a" #i "], %[p]\n\t"
#define SPX_FM(i) "v_fma_f32 %[a" #i "], -%[t" #i "], %[l], %[a" #i "]\n\t"
#define SPX_AO(i) [a##i] "+v"(a[J0 + i])
#define SPX_TO(i) [t##i] "=&s"(t[i])
#define SPX_HEAD "s_nop 1\n\ts_bitcmp0_b32 %[mp], %[c]\n\ts_cbranch_scc1 .Lspx_chunk_skip%=\n\t"
#define SPX_TAIL ".Lspx_chunk_skip%=:\n"
#define SPX_IN [mp] "s"(mp), [c] "i"(C), [p] "s"(p), [l] "v"(l)
template <int J0, int CN, int C, int N>
__device__ __forceinline__ void chunk_update_f32(float (&a)[N], uint32_t mp, int p, float l) {
  static_assert(CN >= 1 && CN <= 8 && J0 + CN <= N, "chunk of 1..8 columns");
  float t[8];
  if constexpr (CN == 1) {
    asm volatile(SPX_HEAD SPX_RL(0) "s_nop 1\n\t" SPX_FM(0) SPX_TAIL
                 : SPX_AO(0), SPX_TO(0) : SPX_IN : "scc");
  } else if constexpr (CN == 2) {
    asm volatile(SPX_HEAD SPX_RL(0) SPX_RL(1) "s_nop 1\n\t" SPX_FM(0) SPX_FM(1) SPX_TAIL
                 : SPX_AO(0), SPX_AO(1), SPX_TO(0), SPX_TO(1) : SPX_IN : "scc");
  } else if constexpr (CN == 3) {
    asm volatile(SPX_HEAD SPX_RL(0) SPX_RL(1) SPX_RL(2) SPX_FM(0) SPX_FM(1) SPX_FM(2) SPX_TAIL
                 : SPX_AO(0), SPX_AO(1), SPX_AO(2), SPX_TO(0), SPX_TO(1), SPX_TO(2)
                 : SPX_IN : "scc");
  } else if constexpr (CN == 4) {
    asm volatile(SPX_HEAD SPX_RL(0) SPX_RL(1) SPX_RL(2) SPX_RL(3) SPX_FM(0) SPX_FM(1) SPX_FM(2)
                     SPX_FM(3) SPX_TAIL
                 : SPX_AO(0), SPX_AO(1), SPX_AO(2), SPX_AO(3), SPX_TO(0), SPX_TO(1), SPX_TO(2),
                   SPX_TO(3)
                 : SPX_IN : "scc");
  } else if constexpr (CN == 5) {
    asm volatile(SPX_HEAD SPX_RL(0) SPX_RL(1) SPX_RL(2) SPX_RL(3) SPX_RL(4) SPX_FM(0) SPX_FM(1)
                     SPX_FM(2) SPX_FM(3) SPX_FM(4) SPX_TAIL
                 : SPX_AO(0), SPX_AO(1), SPX_AO(2), SPX_AO(3), SPX_AO(4), SPX_TO(0), SPX_TO(1),
                   SPX_TO(2), SPX_TO(3), SPX_TO(4)
                 : SPX_IN : "scc");
  } else if constexpr (CN == 6) {
    asm volatile(SPX_HEAD SPX_RL(0) SPX_RL(1) SPX_RL(2) SPX_RL(3) SPX_RL(4) SPX_RL(5) SPX_FM(0)
                     SPX_FM(1) SPX_FM(2) SPX_FM(3) SPX_FM(4) SPX_FM(5) SPX_TAIL
                 : SPX_AO(0), SPX_AO(1), SPX_AO(2), SPX_AO(3), SPX_AO(4), SPX_AO(5), SPX_TO(0),
                   SPX_TO(1), SPX_TO(2), SPX_TO(3), SPX_TO(4), SPX_TO(5)
                 : SPX_IN : "scc");
  } else if constexpr (CN == 7) {
    asm volatile(SPX_HEAD SPX_RL(0) SPX_RL(1) SPX_RL(2) SPX_RL(3) SPX_RL(4) SPX_RL(5) SPX_RL(6)
                     SPX_FM(0) SPX_FM(1) SPX_FM(2) SPX_FM(3) SPX_FM(4) SPX_FM(5) SPX_FM(6)
                         SPX_TAIL
                 : SPX_AO(0), SPX_AO(1), SPX_AO(2), SPX_AO(3), SPX_AO(4), SPX_AO(5), SPX_AO(6),
                   SPX_TO(0), SPX_TO(1), SPX_TO(2), SPX_TO(3), SPX_TO(4), SPX_TO(5), SPX_TO(6)
                 : SPX_IN : "scc");
  } else {
    asm volatile(SPX_HEAD SPX_RL(0) SPX_RL(1) SPX_RL(2) SPX_RL(3) SPX_RL(4) SPX_RL(5) SPX_RL(6)
                     SPX_RL(7) SPX_FM(0) SPX_FM(1) SPX_FM(2) SPX_FM(3) SPX_FM(4) SPX_FM(5)
                         SPX_FM(6) SPX_FM(7) SPX_TAIL
                 : SPX_AO(0), SPX_AO(1), SPX_AO(2), SPX_AO(3), SPX_AO(4), SPX_AO(5), SPX_AO(6),
                   SPX_AO(7), SPX_TO(0), SPX_TO(1), SPX_TO(2), SPX_TO(3), SPX_TO(4), SPX_TO(5),
                   SPX_TO(6), SPX_TO(7)
                 : SPX_IN : "scc");
  }
}
#undef SPX_RL
#undef SPX_FM
#undef SPX_AO
#undef SPX_TO
#undef SPX_HEAD
#undef SPX_TAIL
#undef SPX_IN

// The FP64 form of chunk_update_f32, four columns per asm statement: the
// pivot row's 32-bit halves are broadcast into the fixed SGPR pairs
// s[88:95] (declared clobbered) and each FMA reads its pair.  The halves
// enter as 32-bit operands (the register allocator hands over the
// sub-registers of each row element), since an asm operand cannot name the
// halves of a 64-bit register pair.
#define SPX_RL64(i, s0, s1) "v_readlane_b32 " s0 ", %[lo" #i "], %[p]\n\tv_readlane_b32 " s1 ", %[hi" #i "], %[p]\n\t"
#define SPX_FM64(i, sp) "v_fma_f64 %[a" #i "], -" sp ", %[l], %[a" #i "]\n\t"
#define SPX_AO64(i) [a##i] "+v"(a[J0 + i])
#define SPX_HALVES(i) [lo##i] "v"((uint32_t)(uint64_t)__double_as_longlong(a[J0 + i])), \
                      [hi##i] "v"((uint32_t)((uint64_t)__double_as_longlong(a[J0 + i]) >> 32))
#define SPX_HEAD64 "s_nop 1\n\ts_bitcmp0_b32 %[mp], %[c]\n\ts_cbranch_scc1 .Lspx_chunk64_skip%=\n\t"
#define SPX_TAIL64 ".Lspx_chunk64_skip%=:\n"
#define SPX_IN64 [mp] "s"(mp), [c] "i"(C), [p] "s"(p), [l] "v"(l)
template <int J0, int CN, int C, int N>
__device__ __forceinline__ void chunk_update_f64_part(double (&a)[N], uint32_t mp, int p, double l) {
  static_assert(CN >= 1 && CN <= 4 && J0 + CN <= N, "one to four columns");
  if constexpr (CN == 1) {
    asm volatile(SPX_HEAD64 SPX_RL64(0, "s88", "s89") "s_nop 1\n\t" SPX_FM64(0, "s[88:89]") SPX_TAIL64
                 : SPX_AO64(0) : SPX_HALVES(0), SPX_IN64 : "scc", "s88", "s89");
  } else if constexpr (CN == 2) {
    asm volatile(SPX_HEAD64 SPX_RL64(0, "s88", "s89") SPX_RL64(1, "s90", "s91")
                     SPX_FM64(0, "s[88:89]") SPX_FM64(1, "s[90:91]") SPX_TAIL64
                 : SPX_AO64(0), SPX_AO64(1) : SPX_HALVES(0), SPX_HALVES(1), SPX_IN64
                 : "scc", "s88", "s89", "s90", "s91");
  } else if constexpr (CN == 3) {
    asm volatile(SPX_HEAD64 SPX_RL64(0, "s88", "s89") SPX_RL64(1, "s90", "s91")
                     SPX_RL64(2, "s92", "s93") SPX_FM64(0, "s[88:89]") SPX_FM64(1, "s[90:91]")
                         SPX_FM64(2, "s[92:93]") SPX_TAIL64
                 : SPX_AO64(0), SPX_AO64(1), SPX_AO64(2)
                 : SPX_HALVES(0), SPX_HALVES(1), SPX_HALVES(2), SPX_IN64
                 : "scc", "s88", "s89", "s90", "s91", "s92", "s93");
  } else {
    asm volatile(SPX_HEAD64 SPX_RL64(0, "s88", "s89") SPX_RL64(1, "s90", "s91")
                     SPX_RL64(2, "s92", "s93") SPX_RL64(3, "s94", "s95") SPX_FM64(0, "s[88:89]")
                         SPX_FM64(1, "s[90:91]") SPX_FM64(2, "s[92:93]") SPX_FM64(3, "s[94:95]")
                             SPX_TAIL64
                 : SPX_AO64(0), SPX_AO64(1), SPX_AO64(2), SPX_AO64(3)
                 : SPX_HALVES(0), SPX_HALVES(1), SPX_HALVES(2), SPX_HALVES(3), SPX_IN64
                 : "scc", "s88", "s89", "s90", "s91", "s92", "s93", "s94", "s95");
  }
}
#undef SPX_RL64
#undef SPX_FM64
#undef SPX_AO64
#undef SPX_HALVES
#undef SPX_HEAD64
#undef SPX_TAIL64
#undef SPX_IN64
// A chunk of up to eight FP64 columns: two asm statements of up to four.
template <int J0, int CN, int C, int N>
__device__ __forceinline__ void chunk_update_f64(double (&a)[N], uint32_t mp, int p, double l) {
  if constexpr (CN <= 4) {
    chunk_update_f64_part<J0, CN, C, N>(a, mp, p, l);
  } else {
    chunk_update_f64_part<J0, 4, C, N>(a, mp, p, l);
    chunk_update_f64_part<J0 + 4, CN - 4, C, N>(a, mp, p, l);
  }
}

template <int N, typename T>
struct RowLU {
  T a[N];
  int pos;     // position of this lane's row (== lane once factored)
  int orig;    // row of the input whose factor row lane `lane` holds
  T piv;       // U[lane][lane] once factored
  T dinv;      // 1 / U[lane][lane]
  int ipiv;    // factor<true>: lane k holds the position swapped with k at step k
  bool inexact;  // FP64 fast pass: a near-tie needs factor<*, true> on reloaded rows

  // Factor in place.  Returns false if a pivot column is zero (or NaN): the
  // reference's LU then reports the system singular.  Rows and columns past
  // n are identity padding, whose multipliers are zero.
  //
  // Structural skip: `cmask` holds, per lane, a bit for every chunk of
  // kChunk columns in which this lane's row may be non-zero (a superset is
  // fine; ~0u skips nothing).  Step k reads the pivot row's mask, skips the
  // update of every chunk in which the pivot row is zero — those updates are
  // fma(-0, l, a) = a, so the factors are the same values (up to the sign of
  // a zero, which the pivot search ignores) — and adds the pivot row's
  // chunks to every row that takes a non-zero multiplier (the fill).  Rows
  // of an MNA Jacobian hold a few entries, so most chunks are skipped
  // (a 64-unknown inverter chain: 432 of 2,016 column updates run).
  template <bool Record = false, bool Exact = false>
  __device__ __forceinline__ bool factor(int lane, uint32_t cmask = ~0u) {
    constexpr bool kF64 = std::is_same<T, double>::value;
    pos = lane;
    if constexpr (Record) ipiv = lane;
    dinv = T(0);
    piv = T(0);
    bool ok = true;
    uint32_t nearby = 0;  // FP64 fast pass: a near-tie the exact rule might decide otherwise

    // No early exit: a zero pivot column clears `ok`; its step then runs
    // with a zero reciprocal and the result is discarded by the caller.
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
      if constexpr (k + 1 < N) {
        const uint32_t mp = (uint32_t)__builtin_amdgcn_readlane((int)cmask, p);
        if (l != T(0)) cmask |= mp;
        static_for<(k + 1) / kChunk, (N + kChunk - 1) / kChunk>(
            [&](auto C) __attribute__((always_inline)) {
              constexpr int c = C;
              constexpr int j0 = c * kChunk > k + 1 ? c * kChunk : k + 1;
              constexpr int j1 = (c + 1) * kChunk < N ? (c + 1) * kChunk : N;
              if constexpr (kF64)
                chunk_update_f64<j0, j1 - j0, c>(a, mp, p, l);
              else
                chunk_update_f32<j0, j1 - j0, c>(a, mp, p, l);
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
