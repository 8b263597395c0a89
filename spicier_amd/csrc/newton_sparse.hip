// newton_sparse.hip — Newton DC operating point for systems past the
// register LU (65 <= n <= 4096): one 256-thread workgroup per instance,
// the shared-symbolic LU of sparse_plan.h, FP64 throughout.
//
// Per instance the workgroup keeps x, rhs, the saved stepping solution and
// the triangular-solve vector in LDS; element values, pattern values and the
// filled L+U values live in the workgroup's slot of a global workspace (in
// LDS instead when they fit), which stays in L2 while the instance runs.
// The grid is persistent: each workgroup walks instances blockIdx.x,
// blockIdx.x + gridDim.x, ..., so the workspace scales with the grid, not
// with the batch.
#include <hip/hip_runtime.h>

#include "assembly.h"
#include "dcop_driver.h"
#include "program.h"
#include "sparse_plan.h"

namespace spx {

constexpr int kSparseThreads = 256;

struct SparseInstance {
  static constexpr int T = kSparseThreads;
  const DevProgram& g;
  const DevSparse& sp;
  const DevNrCfg& cfg;
  int tid;
  const double* pv;  // parameters (global, read only)
  double* vals;      // element value slots
  double* vst;       // pnjlim junction states
  double* av;        // assembled pattern values (circuit program order)
  double* lu;        // filled L+U values
  double* rhs;       // LDS n
  double* xs;        // LDS n
  double* xw;        // LDS n
  double* y;         // LDS n, elimination order
  double* lbuf;      // LDS max_s
  double* ubuf;      // LDS max_s
  int* red;          // LDS T/64 ints
  double* redd;      // LDS T/64 doubles
  int nfact = 0;
  int nref = 0;

  __device__ __forceinline__ bool block_all(bool p) {
    const bool w = __all(p);
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = w;
    __syncthreads();
    bool r = true;
    for (int i = 0; i < T / kWave; ++i) r = r && red[i];
    __syncthreads();
    return r;
  }
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
      eval_element(g, g.order[i], xs, pv, vals, vst, sf, limiter, lim);
    if (tid == 0) vals[kValGmin] = gmin;
    __syncthreads();
    for (int q = tid; q < g.num_pos; q += T) {
      const double v = gather_sum(g.pos_off, g.contrib, q, vals);
      av[q] = v;
      lu[sp.pos_lu[q]] = v;
    }
    for (int f = tid; f < sp.num_fill; f += T) lu[sp.fill[f]] = 0.0;
    for (int r = tid; r < g.n; r += T) rhs[r] = gather_sum(g.rhs_off, g.rhs_contrib, r, vals);
    __syncthreads();
    return !block_all(!lim);
  }

  // Factor lu in place and solve for rhs; the solution is left in y in
  // elimination order.  False if a pivot is exactly zero or the solution is
  // not finite.
  __device__ __forceinline__ bool factor_solve() {
    ++nfact;
    const int n = g.n;
    bool ok = true;
    for (int k = 0; k < n; ++k) {
      const int s0 = sp.col_ptr[k], s = sp.col_ptr[k + 1] - s0;
      const double piv = lu[sp.diag[k]];
      ok = ok && piv != 0.0;
      const double inv = piv != 0.0 ? 1.0 / piv : 0.0;
      for (int t = tid; t < s; t += T) {
        const int li = sp.s_l[s0 + t];
        const double l = lu[li] * inv;
        lu[li] = l;
        lbuf[t] = l;
        ubuf[t] = lu[sp.s_u[s0 + t]];
      }
      __syncthreads();
      if (s > 0) {
        const int* tg = sp.tgt + sp.upd_ptr[k];
        const int ss = s * s;
        int a = tid / s, b = tid - (tid / s) * s;
        const int da = T / s, db = T - (T / s) * s;
        for (int p = tid; p < ss; p += T) {
          const int ti = tg[p];
          lu[ti] = fma(-lbuf[a], ubuf[b], lu[ti]);
          a += da;
          b += db;
          if (b >= s) {
            b -= s;
            ++a;
          }
        }
      }
      __syncthreads();
    }
    if (!ok) return false;
    // L y = P b (unit lower, column oriented)
    for (int k = tid; k < n; k += T) y[k] = rhs[sp.perm[k]];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
      const double yk = y[k];
      const int s0 = sp.col_ptr[k], s1 = sp.col_ptr[k + 1];
      for (int t = s0 + tid; t < s1; t += T) {
        const int i = sp.s_row[t];
        y[i] = fma(-lu[sp.s_l[t]], yk, y[i]);
      }
      __syncthreads();
    }
    // U x = y (column oriented, bottom up); y[k] takes x_k one step later,
    // after every thread has read y[k]
    for (int k = n - 1; k >= 0; --k) {
      const double xk = y[k] * (1.0 / lu[sp.diag[k]]);
      const int u0 = sp.ut_ptr[k], u1 = sp.ut_ptr[k + 1];
      for (int t = u0 + tid; t < u1; t += T) {
        const int i = sp.ut_row[t];
        y[i] = fma(-lu[sp.ut_idx[t]], xk, y[i]);
      }
      __syncthreads();
      if (tid == 0) y[k] = xk;
    }
    __syncthreads();
    bool fin = true;
    for (int k = tid; k < n; k += T) fin = fin && isfinite(y[k]);
    return block_all(fin);
  }

  __device__ __forceinline__ int newton(double gmin, double sf, int& iters, int& fp64_escalations) {
    const int n = g.n, nn = g.num_nodes;
    for (int i = tid; i < g.num_elems; i += T) vst[i] = 0.0;
    __syncthreads();
    for (int it = 0; it < cfg.max_iterations; ++it) {
      const bool lim = assemble(gmin, sf, cfg.limiter);
      if (!factor_solve()) {
        iters = it + 1;
        return -1;
      }
      bool pass = true;
      for (int k = tid; k < n; k += T) {  // check_convergence, newton.rs:134-159
        const int u = sp.perm[k];
        const double xn = y[k], xo = xs[u];
        const double delta = fabs(xn - xo);
        const double tol = cfg.v_reltol * fmax(fabs(xn), fabs(xo)) +
                           (u < nn ? cfg.v_abstol : cfg.i_abstol);
        pass = pass && !(delta > tol);
      }
      const bool conv = block_all(pass) && !lim;
      for (int k = tid; k < n; k += T) xs[sp.perm[k]] = y[k];
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
    const bool ok = factor_solve();
    for (int k = tid; k < g.n; k += T) xs[sp.perm[k]] = ok ? y[k] : 0.0;
    __syncthreads();
    return ok ? 0 : 1;
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
  __device__ __forceinline__ void write_handover(const DevOut& out, int64_t inst,
                                                 int flags) const {
    if (tid == 0) {
      out.status[inst] = SPX_ACTIVE;
      if (out.flags) out.flags[inst] = flags | SPX_FLAG_FP64_ESCALATED;
    }
  }
};

// Per-slot global workspace, in doubles.
inline int64_t sparse_slot_doubles(const DevProgram& g, const DevSparse& sp, bool lu_in_lds) {
  int64_t d = (int64_t)g.num_vals + g.num_elems + g.num_pos + (lu_in_lds ? 0 : sp.nnz_lu);
  return (d + 31) & ~int64_t(31);  // 256-byte slots
}

inline size_t sparse_smem_bytes(const DevProgram& g, const DevSparse& sp, bool lu_in_lds) {
  size_t d = 4 * (size_t)g.n + 2 * (size_t)sp.max_s + kSparseThreads / kWave;
  size_t b = d * sizeof(double) + (kSparseThreads / kWave) * sizeof(int);
  b = (b + 7) & ~size_t(7);
  if (lu_in_lds) b += (size_t)sp.nnz_lu * sizeof(double);
  return b;
}

__global__ __launch_bounds__(kSparseThreads) void k_newton_sparse(
    DevProgram g, DevSparse sp, DevNrCfg cfg, const double* __restrict__ params,
    const double* __restrict__ x0, int64_t B, DevOut out, double* ws, int64_t slot,
    int lu_in_lds) {
  extern __shared__ double smem[];
  const int n = g.n, tid = threadIdx.x;
  double* rhs = smem;
  double* xs = rhs + n;
  double* xw = xs + n;
  double* y = xw + n;
  double* lbuf = y + n;
  double* ubuf = lbuf + sp.max_s;
  double* redd = ubuf + sp.max_s;
  int* red = reinterpret_cast<int*>(redd + kSparseThreads / kWave);
  double* lds_lu = reinterpret_cast<double*>(
      (reinterpret_cast<uintptr_t>(red + kSparseThreads / kWave) + 7) & ~uintptr_t(7));
  double* w = ws + blockIdx.x * slot;
  double* vals = w;
  double* vst = vals + g.num_vals;
  double* av = vst + g.num_elems;
  double* lu = lu_in_lds ? lds_lu : av + g.num_pos;
  for (int64_t inst = blockIdx.x; inst < B; inst += gridDim.x) {
    if (tid == 0) {
      vals[kValOne] = 1.0;
      vals[kValGmin] = 0.0;
    }
    for (int i = tid; i < n; i += kSparseThreads) xs[i] = x0 ? x0[inst * n + i] : 0.0;
    __syncthreads();
    SparseInstance I{g, sp, cfg, tid, params + inst * g.P, vals, vst, av, lu, rhs, xs, xw, y,
                     lbuf, ubuf, red, redd};
    run_dcop(I, g, cfg, x0 != nullptr, out, inst, 0);
    __syncthreads();
  }
}

// Grid and workspace shape of a sparse launch: `grid` workgroups, each with
// a slot of `slot` doubles of global workspace and `smem` bytes of LDS.
hipError_t sparse_launch_shape(const DevProgram& g, const DevSparse& sp, int64_t B, int* grid,
                               int64_t* slot, size_t* smem, int* lu_in_lds) {
  const size_t lds_budget = 64 * 1024;
  *lu_in_lds = sparse_smem_bytes(g, sp, true) <= lds_budget ? 1 : 0;
  *smem = sparse_smem_bytes(g, sp, *lu_in_lds);
  *slot = sparse_slot_doubles(g, sp, *lu_in_lds);
  hipError_t e = hipFuncSetAttribute((const void*)k_newton_sparse,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)*smem);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 0, per_cu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_newton_sparse,
                                                        kSparseThreads, *smem)) != hipSuccess)
    return e;
  if (per_cu < 1) per_cu = 1;
  int64_t gsz = (int64_t)cus * per_cu;
  if (gsz > B) gsz = B;
  if (gsz < 1) gsz = 1;
  *grid = (int)gsz;
  return hipSuccess;
}

hipError_t launch_newton_sparse(const DevProgram& g, const DevSparse& sp, const DevNrCfg& cfg,
                                const double* params, const double* x0, int64_t B,
                                const DevOut& out, double* ws, int grid, int64_t slot,
                                size_t smem, int lu_in_lds, hipStream_t st) {
  if (B == 0) return hipSuccess;
  hipLaunchKernelGGL(k_newton_sparse, dim3((unsigned)grid), dim3(kSparseThreads), smem, st, g,
                     sp, cfg, params, x0, B, out, ws, slot, lu_in_lds);
  return hipGetLastError();
}

}  // namespace spx
