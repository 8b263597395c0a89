// newton_sparse.hip — Newton DC operating point for systems past the
// register LU (65 <= n <= 4096): one workgroup per instance, the
// shared-symbolic, level-scheduled LU of sparse_plan.h, in FP64, or in the
// refinement form (FP32 factors, FP64 iterative refinement).
//
// Per instance the workgroup keeps x, rhs and the triangular-solve vector in
// LDS; element values, pattern values, the stepping stages' saved solution
// and the filled L+U values live in the workgroup's slot of a global
// workspace (the L+U values in LDS instead when they fit).
// The grid is persistent: each workgroup walks instances blockIdx.x,
// blockIdx.x + gridDim.x, ..., so the workspace scales with the grid, not
// with the batch.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "assembly.h"
#include "dcop_driver.h"
#include "lu_wave.h"
#include "program.h"
#include "sparse_plan.h"

namespace spx {

constexpr int kSparseThreads = 256;
// The refinement form's workgroup: its FP32 L+U values take most of a CU's
// LDS, so one workgroup per CU, and 16 waves give a level's entries four
// times the load chains 256 threads would (a 32x32 grid with diodes:
// 286 / 237 / 215 ms per 4,096-instance sweep at 256 / 512 / 1,024
// threads, measured with the factors reused across iterations; 160 / 143 ms
// at 512 / 1,024 threads with fresh factors).
constexpr int kSparseIrThreads = 1024;
// Set in a shape's LDS byte count when the plan's tables are staged in LDS
// (LDS sizes stay far below this bit).
constexpr size_t kPlanInLdsBit = size_t(1) << 40;

template <int T>
__device__ __forceinline__ bool block_all_t(bool p, int tid, int* red) {
  const bool w = __all(p);
  if ((tid & (kWave - 1)) == 0) red[tid / kWave] = w;
  __syncthreads();
  bool r = true;
  for (int i = 0; i < T / kWave; ++i) r = r && red[i];
  __syncthreads();
  return r;
}

// v - sum_t val[t] * y[idx[t]] for t = 0 .. len-1, accumulated in that order
// (one fma per term).  The terms' index and value loads (global or LDS) go
// out eight at a time, then their y reads, before the eight fmas: a long row
// of a triangular solve is a few load rounds, not one per four terms.
template <typename V>
__device__ __forceinline__ V row_dot(const V* val, const int* idx, int len, const V* y, V v) {
  constexpr int U = 8;
  int t = 0;
  for (; t + U <= len; t += U) {
    int j[U];
    V a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      j[u] = idx[t + u];
      a[u] = val[t + u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) b[u] = y[j[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) v = fma(-a[u], b[u], v);
  }
#pragma unroll 4
  for (; t < len; ++t) v = fma(-val[t], y[idx[t]], v);
  return v;
}

// Register row width of the dense tail (the plan caps D at this).
constexpr int kTailN = 48;

// The plan's dense tail block (rows and columns n-D .. n-1, D <= kTailN), one
// row per lane of the calling wave: entries the plan does not store are
// zero, lanes past D hold zero rows.  Every index is clamped before its
// load, so no lane reads outside the tables.
template <typename V>
__device__ __forceinline__ void tail_load(const DevSparse& sp, const V* lu, int lane,
                                          V (&a)[kTailN]) {
  // opaque per call: otherwise the compiler hoists the column addresses and
  // tests out of the instance loop as live registers, which spill
  int D = sp.tail;
  asm volatile("" : "+s"(D));
  asm volatile("" : "+v"(lane));
  const int* ti = sp.td_idx + (lane < D ? lane : 0) * D;
  static_for<0, kTailN>([&](auto C) __attribute__((always_inline)) {
    constexpr int c = C;
    const int q0 = ti[c < D ? c : 0];
    const int q = (lane < D && c < D) ? q0 : -1;
    const V v = lu[q >= 0 ? q : 0];
    a[c] = q >= 0 ? v : V(0);
    // eight columns' loads in flight at a time: the scheduler would
    // otherwise issue all 64 index and value loads first and hold them in
    // registers beside the row
    if constexpr ((c & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  });
}

// Factors the dense tail in one wave, after the sparse levels have applied
// every update from earlier columns, and stores the factors back.  Step k:
// the pivot row is lane k; rows below take l = a * (1 / pivot) and the
// update fma(-u, l, a) of each column right of k.  These are the
// multipliers and updates the one-column levels would
// form, in the same order (rows above the pivot take l = 0, which leaves
// them unchanged).  A zero pivot stays in the stored diagonal, where the
// caller's pivot test finds it.
template <typename V>
__device__ __forceinline__ void tail_factor(const DevSparse& sp, V* lu, int lane) {
  // lane opaque, as in tail_load: the 48 lane tests of the steps are not
  // hoisted out of the caller's loops as live masks
  asm volatile("" : "+v"(lane));
  V a[kTailN];
  tail_load(sp, lu, lane, a);
  // steps past D meet zero rows (pivot 0, multipliers 0) and change
  // nothing, so every step runs without a test
  static_for<0, kTailN>([&](auto K) __attribute__((always_inline)) {
    constexpr int k = K;
    const V pv = bcast(a[k], k);
    const V inv = pv != V(0) ? V(1) / pv : V(0);
    const V l = lane > k ? a[k] * inv : V(0);
    if (lane > k) a[k] = l;
    static_for<k + 1, kTailN>([&](auto J) __attribute__((always_inline)) {
      constexpr int j = J;
      a[j] = fma(-bcast(a[j], k), l, a[j]);
      // eight broadcasts (sixteen SGPRs) in flight at a time
      if constexpr ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    });
  });
  int Ds = sp.tail, ls = lane;
  asm volatile("" : "+s"(Ds));
  asm volatile("" : "+v"(ls));
  const int* ti = sp.td_idx + (ls < Ds ? ls : 0) * Ds;
  static_for<0, kTailN>([&](auto C) __attribute__((always_inline)) {
    constexpr int c = C;
    const int q0 = ti[c < Ds ? c : 0];
    if (ls < Ds && c < Ds && q0 >= 0) lu[q0] = a[c];
    if constexpr ((c & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  });
}

// Both triangular solves of the dense tail in one wave: y of the tail rows
// holds b minus their terms in earlier columns on entry and x on return.
// Forward, column k: rows below take fma(-L(r, k), y_k, y_r); backward,
// column k from the last: x_k = y_k * (1 / U(k, k)), rows above take
// fma(-U(r, k), x_k, y_r).
template <typename V>
__device__ __forceinline__ void tail_solve(const DevSparse& sp, const V* lu, V* y, int n,
                                           int lane) {
  int D = sp.tail;  // opaque, as in tail_load
  asm volatile("" : "+s"(D));
  asm volatile("" : "+v"(lane));
  const int t0 = n - D;
  V a[kTailN];
  tail_load(sp, lu, lane, a);
  V v = lane < D ? y[t0 + lane] : V(0);
  const V dinv = V(1) / (lane < D ? lu[sp.diag[t0 + lane]] : V(1));
  // steps past D broadcast a zero and change nothing
  static_for<0, kTailN>([&](auto K) __attribute__((always_inline)) {
    constexpr int k = K;
    const V yk = bcast(v, k);
    if (lane > k) v = fma(-a[k], yk, v);
  });
  static_for_down<0, kTailN>([&](auto K) __attribute__((always_inline)) {
    constexpr int k = K;
    const V xk = bcast(v * dinv, k);
    if (lane < k) v = fma(-a[k], xk, v);
  });
  if (lane < D) y[t0 + lane] = v * dinv;
}

// Runs levels [0, count) of one pass: a run of levels flagged narrow goes
// to thread 0 in order (one barrier after the run), any other level to
// the whole workgroup (barriers inside `wide_fn`).
template <class NarrowFn, class WideFn>
__device__ __forceinline__ void run_levels(int count, const int* flags, int bit, int tid,
                                           NarrowFn narrow_fn, WideFn wide_fn) {
  for (int L = 0; L < count;) {
    if (flags[L] & bit) {
      int L1 = L + 1;
      while (L1 < count && (flags[L1] & bit)) ++L1;
      if (tid == 0)
        for (int l = L; l < L1; ++l) narrow_fn(l);
      __syncthreads();
      L = L1;
    } else {
      wide_fn(L);
      ++L;
    }
  }
}

// A level with no more rows than the workgroup has waves (the narrow end of
// a grid's levels: a few rows of tens of terms each) takes one wave per row
// in the DC kernels: the lanes load a row's terms at once and the wave sums
// them in a fixed tree, one round trip instead of one per eight terms (the
// sum is then ordered by the tree, not by column).
template <int T, bool kTail>
__device__ __forceinline__ bool wave_rows(int r0, int r1) {
  return kTail && T > kWave && r1 - r0 <= T / kWave;
}

// Lanes per row of a DC kernel's solve level: a wave per row when the
// workgroup has a wave for each, else 32 or 16 lanes per row while that
// covers the level in one round; 0 = a thread per row.
template <int T, bool kTail>
__device__ __forceinline__ int lanes_per_row(int r0, int r1) {
  if (!(kTail && T > kWave)) return 0;
  const int rows = r1 - r0;
  return rows <= T / 64 ? 64 : rows <= T / 32 ? 32 : rows <= T / 16 ? 16 : 0;
}

// Sum over aligned groups of g lanes (g = 16, 32 or 64), every lane of a
// group active; for g = 64 the same tree as wave_sum.
template <typename V>
__device__ __forceinline__ V group_sum(V v, int g) {
  for (int off = g >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Factors lu (values of type V: double, or float for the FP32 factors of
// the refinement form) in place with the plan sp.  False if a pivot is
// exactly zero.  Level by level: multipliers of the level's columns, then
// its updates target by target (contributions in column order).  With
// kTail, the plan's dense tail (if any) is factored in wave 0's registers in
// place of its one-column levels.
template <int T, bool kTail, typename V>
__device__ __forceinline__ bool sparse_factor(const DevSparse& sp, int n, V* lu, int tid,
                                              int* red) {
  const int D = kTail ? sp.tail : 0;
  // One level's factor work, threads [t0, t0 + stride).  A thread takes four
  // multipliers, then four targets, at a time and issues their index and
  // value loads together (the values are in L2 or LDS, the plan in L2): a
  // level is a handful of dependent load rounds, not one per entry.  Each
  // target still sums its contributions in column order.  A level with at
  // most one entry per thread (a chain's) takes the plain path, which
  // carries no unused slots.
  auto factor_level = [&](int L, int t0, int stride, bool sync) {
    constexpr int U = 4;
    const int a0 = sp.pa_ptr[L], a1 = sp.pa_ptr[L + 1];
    if (a1 - a0 <= stride) {  // at most one entry per thread (chains): no unrolling
      const int t = a0 + t0;
      if (t < a1) {
        const int li = sp.pa_l[t];
        lu[li] = lu[li] * (V(1) / lu[sp.pa_d[t]]);
      }
    } else
    for (int t = a0 + t0; t < a1; t += U * stride) {
      int li[U], di[U];
      V lv[U], dv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int tt = t + u * stride;
        li[u] = tt < a1 ? sp.pa_l[tt] : -1;
        di[u] = tt < a1 ? sp.pa_d[tt] : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (li[u] >= 0) {
          lv[u] = lu[li[u]];
          dv[u] = lu[di[u]];
        }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (li[u] >= 0) lu[li[u]] = lv[u] * (V(1) / dv[u]);
    }
    if (sync) __syncthreads();
    const int b0 = sp.tb_ptr[L], b1 = sp.tb_ptr[L + 1];
    if (b1 - b0 <= stride) {
      const int t = b0 + t0;
      if (t < b1) {
        const int ti = sp.tb_t[t];
        V v = lu[ti];
        int c = sp.tb_c[t];
        const int c1 = sp.tb_c[t + 1];
        // DC kernels: eight contributions' index loads, then their value
        // loads, then the eight fmas in column order (the transient kernels'
        // register budget has no room for it)
        constexpr int C = 8;
        if constexpr (kTail)
        for (; c + C <= c1; c += C) {
          int cl[C], cu[C];
          V a[C], b[C];
#pragma unroll
          for (int u = 0; u < C; ++u) {
            cl[u] = sp.cb_l[c + u];
            cu[u] = sp.cb_u[c + u];
          }
#pragma unroll
          for (int u = 0; u < C; ++u) {
            a[u] = lu[cl[u]];
            b[u] = lu[cu[u]];
          }
#pragma unroll
          for (int u = 0; u < C; ++u) v = fma(-a[u], b[u], v);
        }
        for (; c < c1; ++c) v = fma(-lu[sp.cb_l[c]], lu[sp.cb_u[c]], v);
        lu[ti] = v;
      }
    } else
    for (int t = b0 + t0; t < b1; t += U * stride) {
      int ti[U], c0[U], c1[U];
      V v[U];
      int cmax = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int tt = t + u * stride;
        const bool in = tt < b1;
        ti[u] = in ? sp.tb_t[tt] : -1;
        c0[u] = in ? sp.tb_c[tt] : 0;
        c1[u] = in ? sp.tb_c[tt + 1] : 0;
        cmax = max(cmax, c1[u] - c0[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (ti[u] >= 0) v[u] = lu[ti[u]];
      // DC kernels: two contributions of each target per round, their index
      // loads, then their value loads, then the fmas in column order
      if constexpr (!kTail) {
        for (int q = 0; q < cmax; ++q) {
          int cl[U], cu[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const bool in = c0[u] + q < c1[u];
            cl[u] = in ? sp.cb_l[c0[u] + q] : -1;
            cu[u] = in ? sp.cb_u[c0[u] + q] : 0;
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (cl[u] >= 0) v[u] = fma(-lu[cl[u]], lu[cu[u]], v[u]);
        }
      } else
      for (int q = 0; q < cmax; q += 2) {
        int cl[2][U], cu[2][U];
        V a[2][U], b[2][U];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const bool in = c0[u] + q + h < c1[u];
            cl[h][u] = in ? sp.cb_l[c0[u] + q + h] : -1;
            cu[h][u] = in ? sp.cb_u[c0[u] + q + h] : 0;
          }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (cl[h][u] >= 0) {
              a[h][u] = lu[cl[h][u]];
              b[h][u] = lu[cu[h][u]];
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (cl[h][u] >= 0) v[u] = fma(-a[h][u], b[h][u], v[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (ti[u] >= 0) lu[ti[u]] = v[u];
    }
    if (sync) __syncthreads();
  };
  run_levels(sp.levels - D, sp.lev_flags, 1, tid, [&](int l) { factor_level(l, 0, 1, false); },
             [&](int l) { factor_level(l, tid, T, true); });
  if (D > 0) {
    if (tid < kWave) tail_factor(sp, lu, tid);
    __syncthreads();
  }
  bool ok = true;
  for (int k = tid; k < n; k += T) ok = ok && lu[sp.diag[k]] != V(0);
  return block_all_t<T>(ok, tid, red);
}

// Four ints of a plan descriptor table (16-byte aligned rows).
__device__ __forceinline__ int4 desc4(const int* p, int r) {
  return *reinterpret_cast<const int4*>(p + 4 * r);
}

// Both triangular solves with the factors of sparse_factor (the same kTail):
// y (LDS, n, elimination order) holds the right-hand side on entry and the
// solution on return.  False if the solution is not finite.  The solves are
// row-oriented over the forward levels and the levels counted from the root.
template <int T, bool kTail, typename V>
__device__ __forceinline__ bool sparse_solve(const DevSparse& sp, int n, const V* lu, V* y,
                                             int tid, int* red) {
  const int D = kTail ? sp.tail : 0;
  // the DC kernels (kTail) take a row's index, L offsets and length from
  // its descriptor in one load; the transient kernels from the row tables
  auto fwd_row = [&](int r) -> int4 {
    if constexpr (kTail) {
      return desc4(sp.fdesc, r);
    } else {
      const int i = sp.lev_rows[r];
      const int u0 = sp.ut_ptr[i], nl = sp.ut_ptr[i + 1] - u0;
      return make_int4(i, u0, nl, sp.diag[i] - nl);
    }
  };
  auto forward_level = [&](int L, int t0, int stride) {
    const int r0 = sp.lev_ptr[L], r1 = sp.lev_ptr[L + 1];
    const int G = stride > 1 ? lanes_per_row<T, kTail>(r0, r1) : 0;
    if (G) {
      const int w = tid / G, lane = tid % G;
      if (r0 + w < r1) {
        const int4 d = fwd_row(r0 + w);
        const V* li = lu + d.w;
        V part = V(0);
        for (int t = lane; t < d.z; t += G) part = fma(li[t], y[sp.ut_row[d.y + t]], part);
        part = group_sum(part, G);
        if (lane == 0) y[d.x] = y[d.x] - part;
      }
      return;
    }
    for (int r = r0 + t0; r < r1; r += stride) {
      const int4 d = fwd_row(r);
      y[d.x] = row_dot(lu + d.w, sp.ut_row + d.y, d.z, y, y[d.x]);
    }
  };
  const int* bptr = D > 0 ? sp.tblev_ptr : sp.blev_ptr;
  const int* brows = D > 0 ? sp.tblev_rows : sp.blev_rows;
  auto bwd_row = [&](int r) -> int4 {
    if constexpr (kTail) {
      return desc4(sp.bkdesc, r);
    } else {
      const int i = brows[r];
      const int s0 = sp.col_ptr[i];
      return make_int4(i, s0, sp.col_ptr[i + 1] - s0, sp.diag[i]);
    }
  };
  auto backward_level = [&](int L, int t0, int stride) {
    const int r0 = bptr[L], r1 = bptr[L + 1];
    const int G = stride > 1 ? lanes_per_row<T, kTail>(r0, r1) : 0;
    if (G) {
      const int w = tid / G, lane = tid % G;
      if (r0 + w < r1) {
        const int4 d = bwd_row(r0 + w);
        const V* ui = lu + d.w;
        V part = V(0);
        for (int t = lane; t < d.z; t += G) part = fma(ui[1 + t], y[sp.s_row[d.y + t]], part);
        part = group_sum(part, G);
        if (lane == 0) y[d.x] = (y[d.x] - part) * (V(1) / ui[0]);
      }
      return;
    }
    for (int r = r0 + t0; r < r1; r += stride) {
      const int4 d = bwd_row(r);
      const V* ui = lu + d.w;
      y[d.x] = row_dot(ui + 1, sp.s_row + d.y, d.z, y, y[d.x]) * (V(1) / ui[0]);
    }
  };
  run_levels(sp.levels - D, sp.lev_flags, 2, tid, [&](int l) { forward_level(l, 0, 1); },
             [&](int l) {
               forward_level(l, tid, T);
               __syncthreads();
             });
  if (D > 0) {
    for (int r = tid; r < D; r += T) {  // the tail rows' terms in earlier columns
      const int i = n - D + r;
      const int u0 = sp.ut_ptr[i], nl = sp.ut_ptr[i + 1] - u0;
      y[i] = row_dot(lu + sp.diag[i] - nl, sp.ut_row + u0, sp.tl_cnt[r], y, y[i]);
    }
    __syncthreads();
    if (tid < kWave) tail_solve(sp, lu, y, n, tid);
    __syncthreads();
  }
  run_levels(D > 0 ? sp.tblevels : sp.blevels, D > 0 ? sp.tblev_flags : sp.blev_flags, 1, tid,
             [&](int l) { backward_level(l, 0, 1); },
             [&](int l) {
               backward_level(l, tid, T);
               __syncthreads();
             });
  bool fin = true;
  for (int k = tid; k < n; k += T) fin = fin && isfinite(y[k]);
  return block_all_t<T>(fin, tid, red);
}

// Factors lu in place with the plan sp, then solves for the right-hand side
// b(i) (i in unknown order); the solution is left in y (LDS, n) in
// elimination order.  False if a pivot is exactly zero or the solution is
// not finite.
template <int T, bool kTail, class RHS>
__device__ __forceinline__ bool sparse_factor_solve(const DevSparse& sp, int n, double* lu, RHS b,
                                                    double* y, int tid, int* red) {
  if (!sparse_factor<T, kTail>(sp, n, lu, tid, red)) return false;
  for (int k = tid; k < n; k += T) y[k] = b(sp.perm[k]);
  __syncthreads();
  return sparse_solve<T, kTail>(sp, n, lu, y, tid, red);
}

// kTail_: the factorisation takes the plan's dense tail in registers (the
// DC kernel; the transient kernels keep the level form and its registers).
// IR_: the refinement form -- FP32 L+U factors of each Newton Jacobian and
// FP64 iterative refinement against the FP64 assembled Jacobian (the
// instance hands itself to the FP64 kernel when FP32 cannot resolve its
// system).
template <int T_, bool kTail_ = false, bool IR_ = false>
struct SparseInstance {
  static constexpr int T = T_;
  static constexpr bool kTail = kTail_;
  static constexpr bool IR = IR_;
  using LV = typename std::conditional<IR_, float, double>::type;
  const DevProgram& g;
  const DevSparse& sp;
  const DevNrCfg& cfg;
  int tid;
  const double* pv;  // parameters (global, read only)
  double* vals;      // element value slots
  double* vst;       // pnjlim junction states
  double* av;        // assembled pattern values (circuit program order)
  LV* lu;            // filled L+U values
  double* rhs;       // LDS n
  double* xs;        // LDS n
  double* xw;        // global slot n: the stepping stages' saved solution
  LV* y;             // LDS n, elimination order
  int* red;          // LDS T/64 ints
  double* redd;      // LDS T/64 doubles
  double* xi = nullptr;  // LDS n, refinement form: the refinement iterate (unknown order)
  int nfact = 0;
  int nref = 0;
  const TranCtx* tc = nullptr;  // transient context (null: DC sweep)

  __device__ __forceinline__ bool block_all(bool p) { return block_all_t<T>(p, tid, red); }
  __device__ __forceinline__ double block_max(double v) {
    const double w = wave_max(v);
    if ((tid & (kWave - 1)) == 0) redd[tid / kWave] = w;
    __syncthreads();
    double r = redd[0];
    for (int i = 1; i < T / kWave; ++i) r = fmax(r, redd[i]);
    __syncthreads();
    return r;
  }

  __device__ __forceinline__ bool assemble(double gmin, double sf, int limiter) {
    bool lim = false;
    for (int i = tid; i < g.num_elems; i += T)
      eval_element(g, g.order[i], xs, pv, vals, vst, sf, limiter, lim, tc);
    if (tid == 0) vals[kValGmin] = gmin;
    __syncthreads();
    for (int q = tid; q < g.num_pos; q += T) {
      const double v = gather_sum(g.pos_off, g.contrib, q, vals);
      av[q] = v;
      if constexpr (!IR) lu[sp.pos_lu[q]] = v;
    }
    if constexpr (!IR)
      for (int f = tid; f < sp.num_fill; f += T) lu[sp.fill[f]] = 0.0;
    for (int r = tid; r < g.n; r += T) rhs[r] = gather_sum(g.rhs_off, g.rhs_contrib, r, vals);
    __syncthreads();
    return !block_all(!lim);
  }

  __device__ __forceinline__ bool factor_solve() {
    ++nfact;
    return sparse_factor_solve<T, kTail>(sp, g.n, lu, [&](int i) { return rhs[i]; }, y, tid, red);
  }

  // Refinement form: the FP32 factors of the assembled Jacobian (fill
  // entries zero).  False if a pivot is exactly zero in FP32.
  __device__ __forceinline__ bool factor32() {
    ++nfact;
    for (int q = tid; q < g.num_pos; q += T) lu[sp.pos_lu[q]] = (LV)av[q];
    for (int f = tid; f < sp.num_fill; f += T) lu[sp.fill[f]] = LV(0);
    __syncthreads();
    return sparse_factor<T, kTail>(sp, g.n, lu, tid, red);
  }

  // Largest of a >= 0 and of b >= 0 over the workgroup, one barrier pair
  // (refinement form: redd holds 2 T/64 doubles).
  __device__ __forceinline__ void block_max2(double& a, double& b) {
    constexpr int W = T / kWave;
    const double wa = wave_max(a), wb = wave_max(b);
    if ((tid & (kWave - 1)) == 0) {
      redd[tid / kWave] = wa;
      redd[W + tid / kWave] = wb;
    }
    __syncthreads();
    double ra = redd[0], rb = redd[W];
    for (int i = 1; i < W; ++i) {
      ra = fmax(ra, redd[i]);
      rb = fmax(rb, redd[W + i]);
    }
    __syncthreads();
    a = ra;
    b = rb;
  }

  // Refinement form: FP64 iterative refinement of J x = rhs with the FP32
  // factors in lu, from x = xs; the iterate is left in xi.  Stops on LAPACK
  // dsgesv's backward-error test
  //   ||rhs - J x||_inf <= tol * ||J||_inf * ||x||_inf,
  //   tol = ir_tol, or sqrt(n) * 2^-53 if that is 0,
  // the test and the contraction rule of the register-LU kernels'
  // refinement.  The residual of row u sums fma(-J(u, c), x_c, .) over the
  // row's pattern entries in column order, from rhs_u.  Returns 0
  // converged, 2 failed (stopped contracting, or not finite).
  __device__ __forceinline__ int refine() {
    const int n = g.n;
    double rs = 0.0;
    for (int u = tid; u < n; u += T) {
      double a = 0.0;
      for (int e = g.row_ptr[u]; e < g.row_ptr[u + 1]; ++e) a += fabs(av[e]);
      rs = fmax(rs, a);
      xi[u] = xs[u];
    }
    const double anrm = block_max(rs);
    const double tol = cfg.ir_tol > 0.0 ? cfg.ir_tol : sqrt((double)n) * 1.1102230246251565e-16;
    double prev = INFINITY;
    for (int s = 0; s <= cfg.ir_max_steps; ++s) {
      double rn = 0.0, xn = 0.0;
      for (int k = tid; k < n; k += T) {  // residual, row perm[k] into y[k]
        const int u = sp.perm[k];
        double r = rhs[u];
        for (int e = g.row_ptr[u]; e < g.row_ptr[u + 1]; ++e)
          r = fma(-av[e], xi[g.pos_rc[e] >> 16], r);
        y[k] = (LV)r;
        rn = fmax(rn, fabs(r));
        xn = fmax(xn, fabs(xi[u]));
      }
      block_max2(rn, xn);
      if (!(rn > tol * anrm * xn)) {  // converged (also catches r == 0)
        bool fin = true;
        for (int u = tid; u < n; u += T) fin = fin && isfinite(xi[u]);
        return block_all(fin) ? 0 : 2;
      }
      if (s == cfg.ir_max_steps) break;
      if (!sparse_solve<T, kTail>(sp, n, lu, y, tid, red)) break;
      ++nref;
      double dn = 0.0;
      for (int k = tid; k < n; k += T) {
        const double d = (double)y[k];
        xi[sp.perm[k]] += d;
        dn = fmax(dn, fabs(d));
      }
      dn = block_max(dn);
      if (!(dn < 0.5 * prev) && s > 0) break;  // stopped contracting
      prev = dn;
    }
    return 2;
  }

  // Refinement form: x = J^{-1} rhs into xi from FP32 factors of this J.
  // The factors are not carried to the next Newton iteration, unlike the
  // register-LU kernels': on a 32x32 grid with diodes, factors one iteration
  // old need twice the refinement steps, which cost more than the sparse
  // refactorisation they save (215 vs 143 ms per 4,096-instance sweep at
  // 1,024 threads).
  // 0 solved, 2 hand over to FP64.
  __device__ __forceinline__ int solve32() {
    if (!factor32()) return 2;
    return refine();
  }

  __device__ __forceinline__ int newton(double gmin, double sf, int& iters, int& fp64_escalations) {
    const int n = g.n, nn = g.num_nodes;
    for (int i = tid; i < g.num_elems; i += T) vst[i] = 0.0;
    __syncthreads();
    for (int it = 0; it < cfg.max_iterations; ++it) {
      const bool lim = assemble(gmin, sf, cfg.limiter);
      bool pass = true;
      if constexpr (IR) {
        if (solve32() != 0) {  // FP32 cannot resolve this system: hand the instance to FP64
          iters = it + 1;
          return -2;
        }
        for (int u = tid; u < n; u += T) {  // check_convergence, newton.rs:134-159
          const double xn = xi[u], xo = xs[u];
          const double delta = fabs(xn - xo);
          const double tol = cfg.v_reltol * fmax(fabs(xn), fabs(xo)) +
                             (u < nn ? cfg.v_abstol : cfg.i_abstol);
          pass = pass && !(delta > tol);
        }
      } else {
        if (!factor_solve()) {
          iters = it + 1;
          return -1;
        }
        for (int k = tid; k < n; k += T) {  // check_convergence, newton.rs:134-159
          const int u = sp.perm[k];
          const double xn = y[k], xo = xs[u];
          const double delta = fabs(xn - xo);
          const double tol = cfg.v_reltol * fmax(fabs(xn), fabs(xo)) +
                             (u < nn ? cfg.v_abstol : cfg.i_abstol);
          pass = pass && !(delta > tol);
        }
      }
      const bool conv = block_all(pass) && !lim;
      if constexpr (IR) {
        for (int u = tid; u < n; u += T) xs[u] = xi[u];
      } else {
        for (int k = tid; k < n; k += T) xs[sp.perm[k]] = y[k];
      }
      __syncthreads();
      if (conv) {
        iters = it + 1;
        return 1;
      }
    }
    iters = cfg.max_iterations;
    return 0;
  }

  __device__ __forceinline__ int linear_solve() {
    assemble(0.0, 1.0, cfg.limiter);
    if constexpr (IR) {
      const int rc = solve32();
      for (int u = tid; u < g.n; u += T) xs[u] = rc == 0 ? xi[u] : 0.0;
      __syncthreads();
      return rc;
    } else {
      const bool ok = factor_solve();
      for (int k = tid; k < g.n; k += T) xs[sp.perm[k]] = ok ? y[k] : 0.0;
      __syncthreads();
      return ok ? 0 : 1;
    }
  }

  __device__ __forceinline__ void save_x() {
    for (int i = tid; i < g.n; i += T) xw[i] = xs[i];
    __syncthreads();
  }
  __device__ __forceinline__ void restore_x() {
    for (int i = tid; i < g.n; i += T) xs[i] = xw[i];
    __syncthreads();
  }
  __device__ __forceinline__ void zero_x() {
    for (int i = tid; i < g.n; i += T) xs[i] = 0.0;
    __syncthreads();
  }

  // max_i |KCL_i| / tol_i at xs with unlimited devices (sparse row products
  // over the assembled pattern)
  __device__ __forceinline__ double kcl(double gmin) {
    assemble(gmin, 1.0, SPX_LIMIT_NONE);
    double q = 0.0;
    for (int r = tid; r < g.n; r += T) {
      double res = -rhs[r], mag = fabs(rhs[r]);
      for (int e = g.row_ptr[r]; e < g.row_ptr[r + 1]; ++e) {
        const double t = av[e] * xs[g.pos_rc[e] >> 16];
        res += t;
        mag = fmax(mag, fabs(t));
      }
      const double tol = cfg.kcl_reltol * mag + (r < g.num_nodes ? cfg.kcl_abstol : cfg.kcl_vntol);
      double qr = fabs(res) / tol;
      if (!(qr == qr)) qr = INFINITY;
      q = fmax(q, qr);
    }
    return block_max(q);
  }

  __device__ __forceinline__ void write_result(const DevOut& out, int64_t inst, int status,
                                               int total, int flags, double res) const {
    const int n = g.n;
    for (int i = tid; i < n; i += T) out.x[inst * n + i] = status == SPX_SINGULAR ? 0.0 : xs[i];
    if (tid == 0) {
      if (out.status) out.status[inst] = status;
      if (out.iterations) out.iterations[inst] = total;
      if (out.flags) out.flags[inst] = flags;
      if (out.residual) out.residual[inst] = res;
      if (out.factorizations) out.factorizations[inst] = nfact;
      if (out.refinements) out.refinements[inst] = nref;
    }
  }
  __device__ __forceinline__ void set_reuse(bool) {}  // FP64 factors: nothing carried over
  // Capacitor and inductor states from xs after a stage (cap_rule's and
  // ind_rule's rules; transient)
  __device__ __forceinline__ void cap_update(int rule, double h) {
    double* vp = const_cast<double*>(tc->vprev);
    double* ip = const_cast<double*>(tc->iprev);
    double* vq = const_cast<double*>(tc->vpp);
    for (int e = tid; e < g.num_elems; e += T) {
      const int* el = g.elems + e * kElemInts;
      if (el[0] == SPX_EL_CAPACITOR) {
        const double v = (el[1] >= 0 ? xs[el[1]] : 0.0) - (el[2] >= 0 ? xs[el[2]] : 0.0);
        cap_rule(rule, pv[el[4]], v, h, vp[e], ip[e], vq[e]);
      } else if (el[0] == SPX_EL_INDUCTOR && ip) {  // el[3]: its branch row
        const double v = (el[1] >= 0 ? xs[el[1]] : 0.0) - (el[2] >= 0 ? xs[el[2]] : 0.0);
        ind_rule(rule, pv[el[4]], v, xs[el[3]], h, vp[e], ip[e], vq[e]);
      }
    }
    __syncthreads();
  }
  // Largest LTE of a trapezoidal step of h over the capacitors and inductors,
  // and the largest capacitor |v| and inductor |i_prev| (adaptive)
  __device__ __forceinline__ void cap_lte_max(double h, double& lte, double& ref) {
    double l = 0.0, r = 0.0;
    for (int e = tid; e < g.num_elems; e += T) {
      const int* el = g.elems + e * kElemInts;
      if (el[0] == SPX_EL_CAPACITOR) {
        const double v = (el[1] >= 0 ? xs[el[1]] : 0.0) - (el[2] >= 0 ? xs[el[2]] : 0.0);
        l = fmax(l, cap_lte(pv[el[4]], v, h, tc->vprev[e], tc->iprev[e]));
        r = fmax(r, fabs(v));
      } else if (el[0] == SPX_EL_INDUCTOR) {  // the reference takes |i_prev| here
        const double v = (el[1] >= 0 ? xs[el[1]] : 0.0) - (el[2] >= 0 ? xs[el[2]] : 0.0);
        l = fmax(l, ind_lte(pv[el[4]], v, h, tc->vprev[e]));
        r = fmax(r, fabs(tc->iprev[e]));
      }
    }
    lte = block_max(l);
    ref = block_max(r);
  }
  __device__ __forceinline__ void write_probes(const DevTran& tr, double* row) const {
    for (int p = tid; p < tr.num_probes; p += T) row[p] = xs[tr.probes[p]];
  }
  template <typename V>
  __device__ __forceinline__ void write_scalar(V* dst, V v) const {
    if (tid == 0) *dst = v;
  }

  __device__ __forceinline__ void write_handover(const DevOut& out, int64_t inst,
                                                 int flags) const {
    if (tid == 0) {
      out.status[inst] = SPX_ACTIVE;
      if (out.flags) out.flags[inst] = flags | SPX_FLAG_FP64_ESCALATED;
    }
  }
};

// Per-slot global workspace, in doubles.
// The stepping stages' saved solution (xw, n) is here too: it is touched
// once per stepping level, and in LDS it would cost a resident workgroup of
// the one-wave transient kernels.  lu_bytes: bytes per L+U value (8, or 4
// for the FP32 factors of the refinement form).
inline int64_t sparse_slot_doubles(const DevProgram& g, const DevSparse& sp, bool lu_in_lds,
                                   int lu_bytes = 8) {
  int64_t d = (int64_t)g.num_vals + g.num_elems + g.num_pos + g.n +
              (lu_in_lds ? 0 : ((int64_t)sp.nnz_lu * lu_bytes + 7) / 8);
  return (d + 31) & ~int64_t(31);  // 256-byte slots
}

// LDS bytes of the DC and transient kernels: rhs, x and the solve vector
// (n doubles each), the reductions, and the L+U values when in LDS.  The
// refinement form (ir) keeps the refinement iterate as its third n doubles,
// the solve vector as n floats, two reduction slots per wave and FP32 L+U.
inline size_t sparse_smem_bytes(const DevProgram& g, const DevSparse& sp, bool lu_in_lds,
                                int threads = kSparseThreads, bool ir = false) {
  const int w = threads / kWave;
  size_t d = 3 * (size_t)g.n + (ir ? 2 : 1) * w;
  size_t b = d * sizeof(double) + w * sizeof(int);
  b = (b + 7) & ~size_t(7);
  if (ir) b = (b + (size_t)g.n * sizeof(float) + 7) & ~size_t(7);
  if (lu_in_lds) b += (size_t)sp.nnz_lu * (ir ? sizeof(float) : sizeof(double));
  return b;
}

// Ints of the DC kernels' staged level tables (stage_level_tables), with
// the alignment pad.
inline size_t sparse_stage_ints(const DevSparse& sp) {
  auto r4 = [](size_t c) { return (c + 3) & ~size_t(3); };
  const size_t bl = sp.tail > 0 ? sp.tblevels : sp.blevels;
  return 4 + 4 * (size_t)sp.n_fdesc + 4 * (size_t)sp.n_bkdesc + 3 * r4(sp.levels + 1) +
         r4(sp.levels) + r4(bl + 1) + r4(bl);
}

// Copies the tables every level of the DC kernels' factorisation and solves
// reads -- the solve descriptors, the level pointers and flags -- to LDS at
// `at` and points `to` at the copies (followed by a barrier).
template <int T>
__device__ __forceinline__ void stage_level_tables(const DevSparse& sp, DevSparse& to,
                                                   const void* at, int tid) {
  int* q = reinterpret_cast<int*>((reinterpret_cast<uintptr_t>(at) + 15) & ~uintptr_t(15));
  auto put = [&](const int* src, int cnt) {
    int* dst = q;
    for (int i = tid; i < cnt; i += T) dst[i] = src[i];
    q += (cnt + 3) & ~3;
    return static_cast<const int*>(dst);
  };
  const bool tl = sp.tail > 0;
  const int bl = tl ? sp.tblevels : sp.blevels;
  to.fdesc = put(sp.fdesc, 4 * sp.n_fdesc);
  to.bkdesc = put(sp.bkdesc, 4 * sp.n_bkdesc);
  to.lev_ptr = put(sp.lev_ptr, sp.levels + 1);
  to.pa_ptr = put(sp.pa_ptr, sp.levels + 1);
  to.tb_ptr = put(sp.tb_ptr, sp.levels + 1);
  to.lev_flags = put(sp.lev_flags, sp.levels);
  const int* bp = put(tl ? sp.tblev_ptr : sp.blev_ptr, bl + 1);
  const int* bf = put(tl ? sp.tblev_flags : sp.blev_flags, bl);
  if (tl) {
    to.tblev_ptr = bp;
    to.tblev_flags = bf;
  } else {
    to.blev_ptr = bp;
    to.blev_flags = bf;
  }
  __syncthreads();
}

// Four waves per SIMD (128 VGPRs): the FP64 kernel then spills the dense
// tail's register row in part (464 B per lane), and four resident
// workgroups per CU instead of three still carry more instances through the
// level chains' round trips (a 32x32 grid with diodes: 122.4 -> 112.6 ms per
// 4,096-instance sweep; five waves spill 624 B and take 155 ms).  The
// 1,024-thread refinement form runs four waves per SIMD in any case.
// IR: the refinement form (FP32 factors, FP64 refinement); it leaves an
// instance FP32 cannot resolve at SPX_ACTIVE.  escalated_only: run only the
// instances left at SPX_ACTIVE (the FP64 pass after the refinement form).
template <int T, bool IR>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(4))) void k_newton_sparse(
    DevProgram g, DevSparse sp, DevNrCfg cfg, const double* __restrict__ params,
    const double* __restrict__ x0, int64_t B, DevOut out, double* ws, int64_t slot,
    int lu_in_lds, int escalated_only, int stage) {
  using LV = typename std::conditional<IR, float, double>::type;
  extern __shared__ double smem[];
  const int n = g.n, tid = threadIdx.x;
  double* rhs = smem;
  double* xs = rhs + n;
  double* x3 = xs + n;  // FP64: the solve vector; refinement form: the iterate
  double* redd = x3 + n;
  int* red = reinterpret_cast<int*>(redd + (IR ? 2 : 1) * (T / kWave));
  LV* after = reinterpret_cast<LV*>(
      (reinterpret_cast<uintptr_t>(red + T / kWave) + 7) & ~uintptr_t(7));
  LV* y;
  LV* lds_lu;
  if constexpr (IR) {
    y = after;
    lds_lu = reinterpret_cast<LV*>((reinterpret_cast<uintptr_t>(after + n) + 7) & ~uintptr_t(7));
  } else {
    y = x3;
    lds_lu = after;
  }
  double* w = ws + blockIdx.x * slot;
  double* vals = w;
  double* vst = vals + g.num_vals;
  double* av = vst + g.num_elems;
  double* xw = av + g.num_pos;
  LV* lu = lu_in_lds ? lds_lu : reinterpret_cast<LV*>(xw + n);
  // stage: the solve descriptors and per-level tables follow in LDS
  DevSparse spl = sp;
  if (stage) stage_level_tables<T>(sp, spl, lds_lu + (lu_in_lds ? sp.nnz_lu : 0), tid);
  for (int64_t inst = blockIdx.x; inst < B; inst += gridDim.x) {
    if (escalated_only && out.status[inst] != SPX_ACTIVE) continue;
    if (tid == 0) {
      vals[kValOne] = 1.0;
      vals[kValGmin] = 0.0;
    }
    for (int i = tid; i < n; i += T) xs[i] = x0 ? x0[inst * n + i] : 0.0;
    __syncthreads();
    SparseInstance<T, true, IR> I{g,  spl, cfg, tid, params + inst * g.P, vals, vst, av, lu,
                                  rhs, xs, xw, y, red, redd, x3};
    run_dcop(I, g, cfg, x0 != nullptr, out, inst, escalated_only);
    __syncthreads();
  }
}

inline hipError_t grid_for(const void* kernel, size_t smem, int64_t B, int* grid,
                           int threads = kSparseThreads) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, smem)) !=
      hipSuccess)
    return e;
  if (per_cu < 1) per_cu = 1;
  int64_t gsz = (int64_t)cus * per_cu;
  if (gsz > B) gsz = B;
  if (gsz < 1) gsz = 1;
  *grid = (int)gsz;
  return hipSuccess;
}

__global__ void k_queue_reset(int* q) { *q = 0; }

// Batched Backward-Euler transient on the sparse path; the workspace slot
// adds the capacitor states after the sparse Newton slot.
// At least three waves per SIMD (the register budget the unrolled level
// loops fit in without spilling; at two, ladders run 1.4x slower).
template <int T, bool GENERAL>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(3))) void k_transient_sparse(
    DevProgram g, DevSparse sp, DevNrCfg cfg, DevTran tr, const double* __restrict__ params,
    int64_t B, DevOut out, DevTranOut tout, double* ws, int64_t slot, int lu_in_lds,
    int plan_in_lds, int* next) {
  extern __shared__ double smem[];
  __shared__ int64_t s_inst;
  const int n = g.n, tid = threadIdx.x;
  double* rhs = smem;
  double* xs = rhs + n;
  double* y = xs + n;
  double* redd = y + n;
  int* red = reinterpret_cast<int*>(redd + T / kWave);
  double* lds_lu = reinterpret_cast<double*>(
      (reinterpret_cast<uintptr_t>(red + T / kWave) + 7) & ~uintptr_t(7));
  double* w = ws + blockIdx.x * slot;
  double* vals = w;
  double* vst = vals + g.num_vals;
  double* av = vst + g.num_elems;
  double* xw = av + g.num_pos;
  double* vcap = xw + n;  // vprev, iprev, vpp: num_elems each
  double* lu = lu_in_lds ? lds_lu : vcap + 3 * g.num_elems;
  // small plans are read from LDS: every elimination level reads its index
  // tables, and from global memory each level would cost L2 round trips
  if (plan_in_lds) {
    int* lp = reinterpret_cast<int*>(lu_in_lds ? lds_lu + sp.nnz_lu : lds_lu);
    for (int i = tid; i < sp.blob_ints; i += T) lp[i] = sp.blob[i];
    sp = rebase_sparse(sp, lp);
    __syncthreads();
  }
  // instances: blockIdx.x, + gridDim.x, ... or, with a queue, the next
  // unclaimed one (a workgroup that starts late takes fewer, so a class that
  // shares the GPU with other launches stays balanced); every workgroup
  // leaves once the instances are exhausted
  for (int64_t inst = blockIdx.x;; inst += gridDim.x) {
    if (next) {
      if (tid == 0) s_inst = atomicAdd(next, 1);
      __syncthreads();
      inst = s_inst;
      __syncthreads();
    }
    if (inst >= B) break;
    if (tid == 0) {
      vals[kValOne] = 1.0;
      vals[kValGmin] = 0.0;
    }
    for (int i = tid; i < n; i += T) xs[i] = 0.0;
    for (int e = tid; e < 3 * g.num_elems; e += T) vcap[e] = 0.0;
    __syncthreads();
    TranCtx tc{0, 0.0, tr.h, vcap, vcap + g.num_elems, vcap + 2 * g.num_elems, 0, 0.0, 0.0, 0.0};
    SparseInstance<T> I{g, sp, cfg, tid, params + inst * g.P, vals, vst, av, lu, rhs, xs, xw, y,
                     red, redd};
    I.tc = &tc;
    run_transient<GENERAL>(I, g, cfg, tr, tc, out, tout, inst);
    __syncthreads();
  }
}

template <int T>
hipError_t sparse_transient_shape_t(const DevProgram& g, const DevSparse& sp, int64_t B,
                                    int* grid, int64_t* slot, size_t* smem, int* lu_in_lds,
                                    int force_lu) {
  const size_t lds_budget = 64 * 1024;
  *lu_in_lds = force_lu >= 0 ? force_lu : (sparse_smem_bytes(g, sp, true, T) <= lds_budget ? 1 : 0);
  if (*lu_in_lds && sparse_smem_bytes(g, sp, true, T) > 160 * 1024) *lu_in_lds = 0;
  *smem = sparse_smem_bytes(g, sp, *lu_in_lds, T);
  // the plan's tables follow in LDS while the whole stays within the budget
  // and costs no resident workgroup: a level then reads its index tables
  // from LDS instead of L2, but a one-wave workgroup is limited by LDS, and
  // fewer instances in flight cost more than the L2 round trips save (a
  // 256-unknown ladder: 3 instead of 8 workgroups per CU, 204 vs 116 ms)
  const size_t plan_bytes = (size_t)sp.blob_ints * sizeof(int);
  const int cap = (int)(*smem > lds_budget ? *smem : lds_budget);
  hipError_t e = hipFuncSetAttribute((const void*)k_transient_sparse<T, false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_transient_sparse<T, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  if (e != hipSuccess) return e;
  if (*smem + plan_bytes <= lds_budget) {
    int without = 0, with = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &without, (const void*)k_transient_sparse<T, false>, T, *smem);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &with, (const void*)k_transient_sparse<T, false>, T, *smem + plan_bytes);
    if (e != hipSuccess) return e;
    if (with >= without) *smem = (*smem + plan_bytes) | kPlanInLdsBit;
  }
  *slot = sparse_slot_doubles(g, sp, *lu_in_lds) + ((3 * g.num_elems + 31) & ~31);
  const size_t bytes = *smem & ~kPlanInLdsBit;
  // sized for the lean kernel; the persistent loop covers any grid
  return grid_for((const void*)k_transient_sparse<T, false>, bytes, B, grid, T);
}

// Shape of a sparse transient launch with `threads` (64 or 256) per
// workgroup; force_lu: -1 = LU values in LDS when they fit in 64 KB,
// 0 / 1 = global / LDS.
hipError_t sparse_transient_shape(const DevProgram& g, const DevSparse& sp, int64_t B, int* grid,
                                  int64_t* slot, size_t* smem, int* lu_in_lds, int threads,
                                  int force_lu) {
  return threads == 64
             ? sparse_transient_shape_t<64>(g, sp, B, grid, slot, smem, lu_in_lds, force_lu)
             : sparse_transient_shape_t<256>(g, sp, B, grid, slot, smem, lu_in_lds, force_lu);
}

hipError_t launch_transient_sparse(const DevProgram& g, const DevSparse& sp, const DevNrCfg& cfg,
                                   const DevTran& tr, const double* params, int64_t B,
                                   const DevOut& out, const DevTranOut& tout, double* ws,
                                   int grid, int64_t slot, size_t smem, int lu_in_lds,
                                   int threads, hipStream_t st, int* next) {
  if (B == 0) return hipSuccess;
  if (next) {  // the queue starts at 0 (a kernel, not a memset: graph-safe)
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, st, next);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int plan = (smem & kPlanInLdsBit) ? 1 : 0;
  smem &= ~kPlanInLdsBit;
  const bool general = tr.adaptive || tr.method != SPX_TRAN_BE || g.has_inductor;
  if (threads == 64 && general)
    hipLaunchKernelGGL((k_transient_sparse<64, true>), dim3((unsigned)grid), dim3(64), smem, st,
                       g, sp, cfg, tr, params, B, out, tout, ws, slot, lu_in_lds, plan, next);
  else if (threads == 64)
    hipLaunchKernelGGL((k_transient_sparse<64, false>), dim3((unsigned)grid), dim3(64), smem, st,
                       g, sp, cfg, tr, params, B, out, tout, ws, slot, lu_in_lds, plan, next);
  else if (general)
    hipLaunchKernelGGL((k_transient_sparse<256, true>), dim3((unsigned)grid), dim3(256), smem,
                       st, g, sp, cfg, tr, params, B, out, tout, ws, slot, lu_in_lds, plan, next);
  else
    hipLaunchKernelGGL((k_transient_sparse<256, false>), dim3((unsigned)grid), dim3(256), smem,
                       st, g, sp, cfg, tr, params, B, out, tout, ws, slot, lu_in_lds, plan, next);
  return hipGetLastError();
}

// Batched sparse solve with a shared pattern (the numeric half of
// FaerSparseCachedBatchedSolver / FaerTripletBatchedSolver,
// crates/spicier-batched-sweep/src/faer_sparse_solver.rs:196-362): system
// `inst` has triplet values values[inst][nnz]; pattern entry q sums the
// triplets pos_trip[pos_off[q] .. pos_off[q+1]) in triplet order.
__global__ __launch_bounds__(kSparseThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_sparse_solve(
    DevSparse sp, int n, int num_pos, const int* __restrict__ pos_off,
    const int* __restrict__ pos_trip, int64_t nnz, const double* __restrict__ values,
    const double* __restrict__ rhs, int64_t B, double* __restrict__ x, uint8_t* singular,
    double* ws, int64_t slot, int lu_in_lds) {
  extern __shared__ double smem[];
  const int tid = threadIdx.x;
  double* y = smem;
  int* red = reinterpret_cast<int*>(y + n);
  double* lu = lu_in_lds ? reinterpret_cast<double*>(
                               (reinterpret_cast<uintptr_t>(red + kSparseThreads / kWave) + 7) &
                               ~uintptr_t(7))
                         : ws + blockIdx.x * slot;
  for (int64_t inst = blockIdx.x; inst < B; inst += gridDim.x) {
    const double* v = values + inst * nnz;
    for (int q = tid; q < num_pos; q += kSparseThreads) {
      double sum = 0.0;
      for (int c = pos_off[q]; c < pos_off[q + 1]; ++c) sum += v[pos_trip[c]];
      lu[sp.pos_lu[q]] = sum;
    }
    for (int f = tid; f < sp.num_fill; f += kSparseThreads) lu[sp.fill[f]] = 0.0;
    __syncthreads();
    const double* b = rhs + inst * n;
    const bool ok = sparse_factor_solve<kSparseThreads, true>(
        sp, n, lu, [&](int i) { return b[i]; }, y, tid, red);
    for (int k = tid; k < n; k += kSparseThreads) x[inst * n + sp.perm[k]] = ok ? y[k] : 0.0;
    if (tid == 0 && singular) singular[inst] = ok ? 0 : 1;
    __syncthreads();
  }
}

// Shape of a batched sparse solve: LU in LDS when it fits in 64 KB.
hipError_t sparse_solve_shape(const DevSparse& sp, int n, int64_t B, int* grid, int64_t* slot,
                              size_t* smem, int* lu_in_lds) {
  size_t base = (size_t)n * sizeof(double) + (kSparseThreads / kWave) * sizeof(int) + 8;
  *lu_in_lds = base + (size_t)sp.nnz_lu * sizeof(double) <= 64 * 1024 ? 1 : 0;
  *smem = base + (*lu_in_lds ? (size_t)sp.nnz_lu * sizeof(double) : 0);
  *slot = *lu_in_lds ? 0 : (((int64_t)sp.nnz_lu + 31) & ~int64_t(31));
  hipError_t e = hipFuncSetAttribute((const void*)k_sparse_solve,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)*smem);
  if (e != hipSuccess) return e;
  return grid_for((const void*)k_sparse_solve, *smem, B, grid);
}

hipError_t launch_sparse_solve(const DevSparse& sp, int n, int num_pos, const int* pos_off,
                               const int* pos_trip, int64_t nnz, const double* values,
                               const double* rhs, int64_t B, double* x, uint8_t* singular,
                               double* ws, int grid, int64_t slot, size_t smem, int lu_in_lds,
                               hipStream_t st) {
  if (B == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sparse_solve, dim3((unsigned)grid), dim3(kSparseThreads), smem, st, sp, n,
                     num_pos, pos_off, pos_trip, nnz, values, rhs, B, x, singular, ws, slot,
                     lu_in_lds);
  return hipGetLastError();
}

// Grid and workspace shape of a sparse DC launch: `grid` workgroups, each
// with a slot of `slot` doubles of global workspace and `smem` bytes of LDS.
// FP64 (ir = 0): the L+U values in LDS when everything fits in 64 KB (more
// resident workgroups beat LDS-resident values past that).  Refinement form
// (ir = 1): the FP32 L+U values in LDS whenever they fit in a workgroup's
// 160 KB -- the factorisation and the refinement's solves then never leave
// LDS -- else in the slot.
hipError_t sparse_launch_shape(const DevProgram& g, const DevSparse& sp, int64_t B, int ir,
                               int* grid, int64_t* slot, size_t* smem, int* lu_in_lds) {
  const int T = ir ? kSparseIrThreads : kSparseThreads;
  const void* kern = ir ? (const void*)k_newton_sparse<kSparseIrThreads, true>
                        : (const void*)k_newton_sparse<kSparseThreads, false>;
  const size_t lds_budget = ir ? 160 * 1024 : 64 * 1024;
  *lu_in_lds = sparse_smem_bytes(g, sp, true, T, ir) <= lds_budget ? 1 : 0;
  *smem = sparse_smem_bytes(g, sp, *lu_in_lds, T, ir);
  *slot = sparse_slot_doubles(g, sp, *lu_in_lds, ir ? 4 : 8);
  hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  // the level tables follow in LDS when they fit and cost no resident
  // workgroup
  const size_t stage_bytes = sparse_stage_ints(sp) * sizeof(int);
  if (*smem + stage_bytes <= 160 * 1024) {
    int without = 0, with = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&without, kern, T, *smem);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&with, kern, T, *smem + stage_bytes);
    if (e != hipSuccess) return e;
    if (with >= without) *smem = (*smem + stage_bytes) | kPlanInLdsBit;
  }
  return grid_for(kern, *smem & ~kPlanInLdsBit, B, grid, T);
}

// ir = 1 runs the refinement form; escalated_only = 1 the FP64 pass over the
// instances a refinement-form launch left at SPX_ACTIVE (needs out.status).
hipError_t launch_newton_sparse(const DevProgram& g, const DevSparse& sp, const DevNrCfg& cfg,
                                const double* params, const double* x0, int64_t B,
                                const DevOut& out, double* ws, int grid, int64_t slot,
                                size_t smem, int lu_in_lds, int ir, int escalated_only,
                                hipStream_t st) {
  if (B == 0) return hipSuccess;
  const int stage = (smem & kPlanInLdsBit) ? 1 : 0;
  smem &= ~kPlanInLdsBit;
  if (ir)
    hipLaunchKernelGGL((k_newton_sparse<kSparseIrThreads, true>), dim3((unsigned)grid),
                       dim3(kSparseIrThreads), smem, st, g, sp, cfg, params, x0, B, out, ws,
                       slot, lu_in_lds, 0, stage);
  else
    hipLaunchKernelGGL((k_newton_sparse<kSparseThreads, false>), dim3((unsigned)grid),
                       dim3(kSparseThreads), smem, st, g, sp, cfg, params, x0, B, out, ws, slot,
                       lu_in_lds, escalated_only, stage);
  return hipGetLastError();
}

}  // namespace spx
