// gmres.hip — batched restarted GMRES with right Jacobi preconditioning, one
// 256-thread workgroup per system, all systems sharing one sparsity pattern.
//
// The iteration is solve_gmres_real_preconditioned
// (crates/spicier-solver/src/gmres/real.rs:208-400) with JacobiPreconditioner
// (crates/spicier-solver/src/preconditioner.rs:43-100): modified Gram-Schmidt
// Arnoldi, Givens rotations on the Hessenberg columns (helpers.rs:40-48), a
// true-residual check at the end of each restart cycle, x starting at x0
// (zero in the reference).  It replaces the Metal sketch's host-driven loop
// (crates/spicier-backend-metal/src/batched_gmres.rs:396-657), which returns
// to the host for every dot product, with one launch that runs every system
// to convergence on the device.
//
// Layout per system: x, the residual / Arnoldi vector w, and the Jacobi
// inverse diagonal in LDS; the assembled CSR values and the Krylov bases V
// ((restart+1) x n) and Z = M^-1 V (restart x n) in the workgroup's slot of a
// global workspace.  The grid is persistent over systems.  Every reduction
// is a fixed-order tree (per-thread partial sums, wave DPP tree, then the
// four waves in order), so results are deterministic.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "gmres.h"

namespace spx {

namespace {

__device__ __forceinline__ double block_sum(double v, double* red, int tid) {
  const double w = wave_sum(v);
  if ((tid & (kWave - 1)) == 0) red[tid / kWave] = w;
  __syncthreads();
  double r = red[0];
  for (int i = 1; i < kGmresThreads / kWave; ++i) r += red[i];
  __syncthreads();
  return r;
}

// y = A v over the CSR positions (row by row, columns ascending)
__device__ __forceinline__ void spmv(const DevGmres& p, const double* av, const double* v,
                                     double* y, int tid) {
  for (int i = tid; i < p.n; i += kGmresThreads) {
    double s = 0.0;
    for (int q = p.row_ptr[i]; q < p.row_ptr[i + 1]; ++q) s += av[q] * v[p.col[q]];
    y[i] = s;
  }
}

}  // namespace

__global__ __launch_bounds__(kGmresThreads) void k_gmres(
    DevGmres p, int64_t nnz, const double* __restrict__ values, const double* __restrict__ rhs,
    const double* __restrict__ x0, int64_t B, int restart, int max_iter, double tol, int jacobi,
    double* __restrict__ x_out, int* iterations, double* residual, uint8_t* converged,
    double* ws, int64_t slot) {
  extern __shared__ double smem[];
  const int n = p.n, tid = threadIdx.x;
  const int m = restart < n ? restart : n;
  double* x = smem;
  double* w = x + n;
  double* inv = w + n;
  double* H = inv + n;              // m columns of m + 1 (column k = H[k*(m+1) ..])
  double* g = H + m * (m + 1);      // m + 1
  double* cs = g + m + 1;           // m
  double* sn = cs + m;              // m
  double* y = sn + m;               // m
  double* red = y + m;              // kGmresThreads / kWave
  int* flag = reinterpret_cast<int*>(red + kGmresThreads / kWave);
  double* slotp = ws + blockIdx.x * slot;
  double* av = slotp;
  double* V = av + ((p.num_pos + 31) & ~31);
  double* Z = V + (int64_t)(m + 1) * n;

  for (int64_t inst = blockIdx.x; inst < B; inst += gridDim.x) {
    const double* vals = values + inst * nnz;
    const double* b = rhs + inst * n;
    for (int q = tid; q < p.num_pos; q += kGmresThreads) {
      double s = 0.0;
      for (int c = p.pos_off[q]; c < p.pos_off[q + 1]; ++c) s += vals[p.pos_trip[c]];
      av[q] = s;
    }
    __syncthreads();
    double part = 0.0;
    for (int i = tid; i < n; i += kGmresThreads) {
      const int dp = p.diag_pos[i];
      const double d = dp >= 0 ? av[dp] : 0.0;
      inv[i] = (!jacobi || fabs(d) < 1e-30) ? 1.0 : 1.0 / d;
      x[i] = x0 ? x0[inst * n + i] : 0.0;
      part += b[i] * b[i];
    }
    const double bn = sqrt(block_sum(part, red, tid));
    int total = 0, conv = 0;
    double rel = 0.0;
    if (bn < 1e-30) {  // real.rs:228-236
      for (int i = tid; i < n; i += kGmresThreads) x[i] = 0.0;
      conv = 1;
    } else {
      for (int cycle = 0; cycle < max_iter; ++cycle) {
        // r = b - A x
        spmv(p, av, x, w, tid);
        __syncthreads();
        part = 0.0;
        for (int i = tid; i < n; i += kGmresThreads) {
          w[i] = b[i] - w[i];
          part += w[i] * w[i];
        }
        const double rn = sqrt(block_sum(part, red, tid));
        if (rn / bn < tol) {
          rel = rn / bn;
          conv = 1;
          break;
        }
        const double ir = 1.0 / rn;
        for (int i = tid; i < n; i += kGmresThreads) V[i] = w[i] * ir;
        if (tid == 0) {
          g[0] = rn;
          for (int j = 1; j <= m; ++j) g[j] = 0.0;
        }
        __syncthreads();
        int k = 0;
        while (k < m) {
          ++total;
          if (total > max_iter) break;
          const double* vk = V + (int64_t)k * n;
          double* zk = Z + (int64_t)k * n;
          for (int i = tid; i < n; i += kGmresThreads) zk[i] = vk[i] * inv[i];
          __syncthreads();
          spmv(p, av, zk, w, tid);
          __syncthreads();
          double* hk = H + k * (m + 1);
          for (int j = 0; j <= k; ++j) {  // modified Gram-Schmidt
            const double* vj = V + (int64_t)j * n;
            part = 0.0;
            for (int i = tid; i < n; i += kGmresThreads) part += vj[i] * w[i];
            const double hij = block_sum(part, red, tid);
            if (tid == 0) hk[j] = hij;
            for (int i = tid; i < n; i += kGmresThreads) w[i] -= hij * vj[i];
          }
          part = 0.0;
          for (int i = tid; i < n; i += kGmresThreads) part += w[i] * w[i];
          const double wn = sqrt(block_sum(part, red, tid));
          if (tid == 0) hk[k + 1] = wn;
          if (wn < 1e-30) {  // lucky breakdown
            ++k;
            __syncthreads();
            break;
          }
          double* vk1 = V + (int64_t)(k + 1) * n;
          const double iw = 1.0 / wn;
          for (int i = tid; i < n; i += kGmresThreads) vk1[i] = w[i] * iw;
          if (tid == 0) {
            for (int j = 0; j < k; ++j) {
              const double t = cs[j] * hk[j] + sn[j] * hk[j + 1];
              hk[j + 1] = -sn[j] * hk[j] + cs[j] * hk[j + 1];
              hk[j] = t;
            }
            double c = 1.0, s = 0.0;
            if (!(fabs(hk[k + 1]) < 1e-30)) {
              const double rr = sqrt(hk[k] * hk[k] + hk[k + 1] * hk[k + 1]);
              c = hk[k] / rr;
              s = hk[k + 1] / rr;
            }
            cs[k] = c;
            sn[k] = s;
            hk[k] = c * hk[k] + s * hk[k + 1];
            hk[k + 1] = 0.0;
            const double tg = c * g[k] + s * g[k + 1];
            g[k + 1] = -s * g[k] + c * g[k + 1];
            g[k] = tg;
            *flag = fabs(g[k + 1]) / bn < tol;
          }
          __syncthreads();
          const int done = *flag;
          ++k;
          __syncthreads();
          if (done) break;
        }
        if (tid == 0) {  // back substitution H y = g
          for (int i = k - 1; i >= 0; --i) {
            double sum = g[i];
            for (int j = i + 1; j < k; ++j) sum -= H[j * (m + 1) + i] * y[j];
            y[i] = fabs(H[i * (m + 1) + i]) > 1e-30 ? sum / H[i * (m + 1) + i] : 0.0;
          }
        }
        __syncthreads();
        for (int i = tid; i < n; i += kGmresThreads) {
          double xi = x[i];
          for (int j = 0; j < k; ++j) xi += Z[(int64_t)j * n + i] * y[j];
          x[i] = xi;
        }
        __syncthreads();
        spmv(p, av, x, w, tid);
        __syncthreads();
        part = 0.0;
        for (int i = tid; i < n; i += kGmresThreads) {
          const double d = b[i] - w[i];
          part += d * d;
        }
        const double fr = sqrt(block_sum(part, red, tid));
        rel = fr / bn;
        if (rel < tol || total >= max_iter) {
          conv = rel < tol;
          break;
        }
      }
    }
    for (int i = tid; i < n; i += kGmresThreads) x_out[inst * n + i] = x[i];
    if (tid == 0) {
      if (iterations) iterations[inst] = total;
      if (residual) residual[inst] = rel;
      if (converged) converged[inst] = (uint8_t)conv;
    }
    __syncthreads();
  }
}

size_t gmres_smem_bytes(int n, int restart) {
  const int m = restart < n ? restart : n;
  return sizeof(double) * ((size_t)3 * n + (size_t)m * (m + 1) + (m + 1) + 3 * m +
                           kGmresThreads / kWave) + 16;
}

int64_t gmres_slot_doubles(int n, int num_pos, int restart) {
  const int m = restart < n ? restart : n;
  return (((int64_t)num_pos + 31) & ~int64_t(31)) + (int64_t)(2 * m + 1) * n + 32;
}

hipError_t gmres_grid(int n, int restart, int64_t B, int* grid) {
  const size_t smem = gmres_smem_bytes(n, restart);
  hipError_t e = hipFuncSetAttribute((const void*)k_gmres,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 0, per_cu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_gmres,
                                                        kGmresThreads, smem)) != hipSuccess)
    return e;
  if (per_cu < 1) per_cu = 1;
  int64_t gsz = (int64_t)cus * per_cu;
  if (gsz > B) gsz = B;
  if (gsz < 1) gsz = 1;
  *grid = (int)gsz;
  return hipSuccess;
}

hipError_t launch_gmres(const DevGmres& p, int64_t nnz, const double* values, const double* rhs,
                        const double* x0, int64_t B, int restart, int max_iter, double tol,
                        int jacobi, double* x, int* iters, double* res, uint8_t* conv,
                        double* ws, int grid, hipStream_t st) {
  if (B == 0) return hipSuccess;
  const size_t smem = gmres_smem_bytes(p.n, restart);
  const int64_t slot = gmres_slot_doubles(p.n, p.num_pos, restart);
  hipLaunchKernelGGL(k_gmres, dim3((unsigned)grid), dim3(kGmresThreads), smem, st, p, nnz, values,
                     rhs, x0, B, restart, max_iter, tol, jacobi, x, iters, res, conv, ws, slot);
  return hipGetLastError();
}

}  // namespace spx
