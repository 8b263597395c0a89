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
// Two forms share the Krylov solver (gmres_run<T>):
//   * k_gmres: everything in FP64;
//   * k_gmres_ir: FP64 iterative refinement around an FP32 GMRES — the
//     residual b - A x and the update of x in FP64, the correction A d = r
//     solved with FP32 Krylov vectors, inner products and Givens rotations
//     (half the bytes per Arnoldi step), to a relative 1e-4 per refinement
//     step or to what the outer tolerance still needs.
//
// Layout per system: x, the residual / Arnoldi vector w, and the Jacobi
// inverse diagonal in LDS; the assembled CSR values and the Krylov bases V
// ((restart+1) x n) and Z = M^-1 V (restart x n) in the workgroup's slot of a
// global workspace.  The grid is persistent over systems.  Every reduction
// is a fixed-order tree (per-thread partial sums, wave tree, then the four
// waves in order), so results are deterministic.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "gmres.h"

namespace spx {

namespace {

__device__ __forceinline__ float wave_sum_t(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ double wave_sum_t(double v) { return wave_sum(v); }

// Elements of w and v_j per thread held in registers by the Arnoldi step
// (systems up to kGmresThreads * kMgsRegs unknowns).
constexpr int kMgsRegs = 8;

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red, int tid) {
  const T w = wave_sum_t(v);
  if ((tid & (kWave - 1)) == 0) red[tid / kWave] = w;
  __syncthreads();
  T r = red[0];
  for (int i = 1; i < kGmresThreads / kWave; ++i) r += red[i];
  __syncthreads();
  return r;
}

// y = A v over the CSR positions (row by row, columns ascending)
template <typename T>
__device__ __forceinline__ void spmv(const DevGmres& p, const T* av, const T* v, T* y, int tid) {
  for (int i = tid; i < p.n; i += kGmresThreads) {
    T s = T(0);
    for (int q = p.row_ptr[i]; q < p.row_ptr[i + 1]; ++q) s += av[q] * v[p.col[q]];
    y[i] = s;
  }
}

// Small per-system state in LDS, in the Krylov arithmetic type T.
template <typename T>
struct GmresState {
  T* w;     // n: residual / Arnoldi vector
  T* inv;   // n: Jacobi inverse diagonal
  T* H;     // m columns of m + 1 (column k = H[k*(m+1) ..])
  T* g;     // m + 1
  T* cs;    // m
  T* sn;    // m
  T* y;     // m
  T* red;   // kGmresThreads / kWave
  int* flag;
  T* V;     // (m + 1) x n, global slot
  T* Z;     // m x n, global slot
};

// Restarted GMRES on A x = b from the x in LDS, all in T, until the true
// relative residual |b - A x| / bn is below tol or max_iter Arnoldi steps are
// spent.  Returns the steps taken; rel and conv report the final state.
template <typename T>
__device__ int gmres_run(const DevGmres& p, const T* av, const T* b, T* x, const GmresState<T>& s,
                         int m, int max_iter, T tol, T bn, int tid, T& rel, int& conv) {
  const int n = p.n;
  int total = 0;
  conv = 0;
  rel = T(0);
  for (int cycle = 0; cycle < max_iter; ++cycle) {
    // r = b - A x
    spmv(p, av, x, s.w, tid);
    __syncthreads();
    T part = T(0);
    for (int i = tid; i < n; i += kGmresThreads) {
      s.w[i] = b[i] - s.w[i];
      part += s.w[i] * s.w[i];
    }
    const T rn = sqrt(block_sum(part, s.red, tid));
    if (rn / bn < tol) {
      rel = rn / bn;
      conv = 1;
      break;
    }
    const T ir = T(1) / rn;
    for (int i = tid; i < n; i += kGmresThreads) s.V[i] = s.w[i] * ir;
    if (tid == 0) {
      s.g[0] = rn;
      for (int j = 1; j <= m; ++j) s.g[j] = T(0);
    }
    __syncthreads();
    int k = 0;
    while (k < m) {
      ++total;
      if (total > max_iter) break;
      const T* vk = s.V + (int64_t)k * n;
      T* zk = s.Z + (int64_t)k * n;
      for (int i = tid; i < n; i += kGmresThreads) zk[i] = vk[i] * s.inv[i];
      __syncthreads();
      spmv(p, av, zk, s.w, tid);
      __syncthreads();
      T* hk = s.H + k * (m + 1);
      T* vk1 = s.V + (int64_t)(k + 1) * n;
      T wn;
      if (n <= kGmresThreads * kMgsRegs) {
        // modified Gram-Schmidt with this thread's elements of w and of
        // v_j in registers: v_j is read once (the projection reuses it) and
        // v_{j+1} is loaded while h_j is reduced.  The same operations in
        // the same order as the loop below.
        T wr[kMgsRegs], vr[kMgsRegs];
#pragma unroll
        for (int r = 0; r < kMgsRegs; ++r) {
          const int i = tid + r * kGmresThreads;
          wr[r] = i < n ? s.w[i] : T(0);
          vr[r] = i < n ? s.V[i] : T(0);
        }
        for (int j = 0; j <= k; ++j) {
          T vn[kMgsRegs];
          const T* vj1 = s.V + (int64_t)(j < k ? j + 1 : j) * n;
#pragma unroll
          for (int r = 0; r < kMgsRegs; ++r) {
            const int i = tid + r * kGmresThreads;
            vn[r] = i < n ? vj1[i] : T(0);
          }
          part = T(0);
#pragma unroll
          for (int r = 0; r < kMgsRegs; ++r)
            if (tid + r * kGmresThreads < n) part += vr[r] * wr[r];
          const T hij = block_sum(part, s.red, tid);
          if (tid == 0) hk[j] = hij;
#pragma unroll
          for (int r = 0; r < kMgsRegs; ++r) {
            wr[r] -= hij * vr[r];
            vr[r] = vn[r];
          }
        }
        part = T(0);
#pragma unroll
        for (int r = 0; r < kMgsRegs; ++r)
          if (tid + r * kGmresThreads < n) part += wr[r] * wr[r];
        wn = sqrt(block_sum(part, s.red, tid));
        if (tid == 0) hk[k + 1] = wn;
        if (wn < T(1e-30)) {  // lucky breakdown
          ++k;
          __syncthreads();
          break;
        }
        const T iw = T(1) / wn;
#pragma unroll
        for (int r = 0; r < kMgsRegs; ++r) {
          const int i = tid + r * kGmresThreads;
          if (i < n) vk1[i] = wr[r] * iw;
        }
      } else {
        for (int j = 0; j <= k; ++j) {  // modified Gram-Schmidt
          const T* vj = s.V + (int64_t)j * n;
          part = T(0);
          for (int i = tid; i < n; i += kGmresThreads) part += vj[i] * s.w[i];
          const T hij = block_sum(part, s.red, tid);
          if (tid == 0) hk[j] = hij;
          for (int i = tid; i < n; i += kGmresThreads) s.w[i] -= hij * vj[i];
        }
        part = T(0);
        for (int i = tid; i < n; i += kGmresThreads) part += s.w[i] * s.w[i];
        wn = sqrt(block_sum(part, s.red, tid));
        if (tid == 0) hk[k + 1] = wn;
        if (wn < T(1e-30)) {  // lucky breakdown
          ++k;
          __syncthreads();
          break;
        }
        const T iw = T(1) / wn;
        for (int i = tid; i < n; i += kGmresThreads) vk1[i] = s.w[i] * iw;
      }
      if (tid == 0) {
        for (int j = 0; j < k; ++j) {
          const T t = s.cs[j] * hk[j] + s.sn[j] * hk[j + 1];
          hk[j + 1] = -s.sn[j] * hk[j] + s.cs[j] * hk[j + 1];
          hk[j] = t;
        }
        T c = T(1), sv = T(0);
        if (!(fabs(hk[k + 1]) < T(1e-30))) {
          const T rr = sqrt(hk[k] * hk[k] + hk[k + 1] * hk[k + 1]);
          c = hk[k] / rr;
          sv = hk[k + 1] / rr;
        }
        s.cs[k] = c;
        s.sn[k] = sv;
        hk[k] = c * hk[k] + sv * hk[k + 1];
        hk[k + 1] = T(0);
        const T tg = c * s.g[k] + sv * s.g[k + 1];
        s.g[k + 1] = -sv * s.g[k] + c * s.g[k + 1];
        s.g[k] = tg;
        *s.flag = fabs(s.g[k + 1]) / bn < tol;
      }
      __syncthreads();
      const int done = *s.flag;
      ++k;
      __syncthreads();
      if (done) break;
    }
    if (tid == 0) {  // back substitution H y = g
      for (int i = k - 1; i >= 0; --i) {
        T sum = s.g[i];
        for (int j = i + 1; j < k; ++j) sum -= s.H[j * (m + 1) + i] * s.y[j];
        s.y[i] = fabs(s.H[i * (m + 1) + i]) > T(1e-30) ? sum / s.H[i * (m + 1) + i] : T(0);
      }
    }
    __syncthreads();
    for (int i = tid; i < n; i += kGmresThreads) {
      T xi = x[i];
      for (int j = 0; j < k; ++j) xi += s.Z[(int64_t)j * n + i] * s.y[j];
      x[i] = xi;
    }
    __syncthreads();
    spmv(p, av, x, s.w, tid);
    __syncthreads();
    part = T(0);
    for (int i = tid; i < n; i += kGmresThreads) {
      const T d = b[i] - s.w[i];
      part += d * d;
    }
    const T fr = sqrt(block_sum(part, s.red, tid));
    rel = fr / bn;
    if (rel < tol || total >= max_iter) {
      conv = rel < tol;
      break;
    }
  }
  return total;
}

// LDS of the small per-system state of gmres_run<T>, after `base` bytes.
template <typename T>
__device__ __forceinline__ GmresState<T> carve(unsigned char* base, int n, int m, T* V, T* Z) {
  T* t = reinterpret_cast<T*>(base);
  GmresState<T> s;
  s.w = t;
  s.inv = s.w + n;
  s.H = s.inv + n;
  s.g = s.H + m * (m + 1);
  s.cs = s.g + m + 1;
  s.sn = s.cs + m;
  s.y = s.sn + m;
  s.red = s.y + m;
  s.flag = reinterpret_cast<int*>(s.red + kGmresThreads / kWave);
  s.V = V;
  s.Z = Z;
  return s;
}

template <typename T>
constexpr size_t state_bytes(int n, int m) {
  return sizeof(T) * ((size_t)2 * n + (size_t)m * (m + 1) + (m + 1) + 3 * m + kGmresThreads / kWave) +
         16;
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
  double* slotp = ws + blockIdx.x * slot;
  double* av = slotp;
  double* V = av + ((p.num_pos + 31) & ~31);
  double* Z = V + (int64_t)(m + 1) * n;
  const GmresState<double> s =
      carve<double>(reinterpret_cast<unsigned char*>(x + n), n, m, V, Z);

  for (int64_t inst = blockIdx.x; inst < B; inst += gridDim.x) {
    const double* vals = values + inst * nnz;
    const double* b = rhs + inst * n;
    for (int q = tid; q < p.num_pos; q += kGmresThreads) {
      double sum = 0.0;
      for (int c = p.pos_off[q]; c < p.pos_off[q + 1]; ++c) sum += vals[p.pos_trip[c]];
      av[q] = sum;
    }
    __syncthreads();
    double part = 0.0;
    for (int i = tid; i < n; i += kGmresThreads) {
      const int dp = p.diag_pos[i];
      const double d = dp >= 0 ? av[dp] : 0.0;
      s.inv[i] = (!jacobi || fabs(d) < 1e-30) ? 1.0 : 1.0 / d;
      x[i] = x0 ? x0[inst * n + i] : 0.0;
      part += b[i] * b[i];
    }
    const double bn = sqrt(block_sum(part, s.red, tid));
    int total = 0, conv = 0;
    double rel = 0.0;
    if (bn < 1e-30) {  // real.rs:228-236
      for (int i = tid; i < n; i += kGmresThreads) x[i] = 0.0;
      conv = 1;
    } else {
      total = gmres_run<double>(p, av, b, x, s, m, max_iter, tol, bn, tid, rel, conv);
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

// FP64 refinement around FP32 GMRES.  LDS: x, r (FP64), then the FP32 right
// hand side, correction and Krylov state.  Slot: FP64 pattern values, their
// FP32 copy, FP32 V and Z.
constexpr int kIrMaxOuter = 30;
constexpr float kIrInnerTol = 1e-4f;

__global__ __launch_bounds__(kGmresThreads) void k_gmres_ir(
    DevGmres p, int64_t nnz, const double* __restrict__ values, const double* __restrict__ rhs,
    const double* __restrict__ x0, int64_t B, int restart, int max_iter, double tol, int jacobi,
    double* __restrict__ x_out, int* iterations, double* residual, uint8_t* converged,
    double* ws, int64_t slot) {
  extern __shared__ double smem[];
  const int n = p.n, tid = threadIdx.x;
  const int m = restart < n ? restart : n;
  double* red = smem;  // FP64 reductions, kGmresThreads / kWave
  double* x = red + kGmresThreads / kWave;
  double* r = x + n;
  float* r32 = reinterpret_cast<float*>(r + n);
  float* d32 = r32 + n;
  double* slotp = ws + blockIdx.x * slot;
  double* av = slotp;
  float* av32 = reinterpret_cast<float*>(av + ((p.num_pos + 31) & ~31));
  float* V = av32 + ((p.num_pos + 63) & ~63);
  float* Z = V + (int64_t)(m + 1) * n;
  const GmresState<float> s = carve<float>(reinterpret_cast<unsigned char*>(d32 + n), n, m, V, Z);

  for (int64_t inst = blockIdx.x; inst < B; inst += gridDim.x) {
    const double* vals = values + inst * nnz;
    const double* b = rhs + inst * n;
    for (int q = tid; q < p.num_pos; q += kGmresThreads) {
      double sum = 0.0;
      for (int c = p.pos_off[q]; c < p.pos_off[q + 1]; ++c) sum += vals[p.pos_trip[c]];
      av[q] = sum;
      av32[q] = (float)sum;
    }
    __syncthreads();
    double part = 0.0;
    for (int i = tid; i < n; i += kGmresThreads) {
      const int dp = p.diag_pos[i];
      const double d = dp >= 0 ? av[dp] : 0.0;
      s.inv[i] = (!jacobi || fabs(d) < 1e-30) ? 1.0f : (float)(1.0 / d);
      x[i] = x0 ? x0[inst * n + i] : 0.0;
      part += b[i] * b[i];
    }
    const double bn = sqrt(block_sum(part, red, tid));
    int total = 0, conv = 0;
    double rel = 0.0;
    if (bn < 1e-30) {
      for (int i = tid; i < n; i += kGmresThreads) x[i] = 0.0;
      conv = 1;
    } else {
      for (int outer = 0; outer < kIrMaxOuter; ++outer) {
        spmv(p, av, x, r, tid);  // r = b - A x in FP64
        __syncthreads();
        part = 0.0;
        for (int i = tid; i < n; i += kGmresThreads) {
          const double ri = b[i] - r[i];
          r[i] = ri;
          part += ri * ri;
        }
        const double rn = sqrt(block_sum(part, red, tid));
        rel = rn / bn;
        if (rel < tol) {
          conv = 1;
          break;
        }
        if (total >= max_iter) break;
        // correction A d = r in FP32, on r scaled to unit norm (FP32 range)
        const double sc = 1.0 / rn;
        part = 0.0;
        for (int i = tid; i < n; i += kGmresThreads) {
          r32[i] = (float)(r[i] * sc);
          d32[i] = 0.0f;
          part += (double)r32[i] * (double)r32[i];
        }
        const float rn32 = (float)sqrt(block_sum(part, red, tid));
        const float itol = (float)fmax((double)kIrInnerTol, tol * bn / rn);
        float irel = 0.0f;
        int iconv = 0;
        const int steps = gmres_run<float>(p, av32, r32, d32, s, m, max_iter - total, itol, rn32,
                                           tid, irel, iconv);
        total += steps;
        for (int i = tid; i < n; i += kGmresThreads) x[i] += (double)d32[i] / sc;
        __syncthreads();
        if (steps == 0 || !(irel < 1.0f)) break;  // the FP32 solve made no progress
      }
      if (!conv) {  // the residual after the last correction
        spmv(p, av, x, r, tid);
        __syncthreads();
        part = 0.0;
        for (int i = tid; i < n; i += kGmresThreads) part += (b[i] - r[i]) * (b[i] - r[i]);
        rel = sqrt(block_sum(part, red, tid)) / bn;
        conv = rel < tol;
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

size_t gmres_smem_bytes(int n, int restart, int precision) {
  const int m = restart < n ? restart : n;
  if (precision == 0) return sizeof(double) * (size_t)n + state_bytes<double>(n, m);
  return sizeof(double) * (kGmresThreads / kWave + 2 * (size_t)n) +
         sizeof(float) * 2 * (size_t)n + state_bytes<float>(n, m);
}

int64_t gmres_slot_doubles(int n, int num_pos, int restart, int precision) {
  const int m = restart < n ? restart : n;
  if (precision == 0)
    return (((int64_t)num_pos + 31) & ~int64_t(31)) + (int64_t)(2 * m + 1) * n + 32;
  return (((int64_t)num_pos + 31) & ~int64_t(31)) + (((int64_t)num_pos + 63) & ~int64_t(63)) / 2 +
         ((int64_t)(2 * m + 1) * n + 1) / 2 + 32;
}

hipError_t gmres_grid(int n, int restart, int precision, int64_t B, int* grid) {
  const size_t smem = gmres_smem_bytes(n, restart, precision);
  const void* kern = precision == 0 ? (const void*)k_gmres : (const void*)k_gmres_ir;
  hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 0, per_cu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kGmresThreads, smem)) !=
      hipSuccess)
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
                        int jacobi, int precision, double* x, int* iters, double* res,
                        uint8_t* conv, double* ws, int grid, hipStream_t st) {
  if (B == 0) return hipSuccess;
  const size_t smem = gmres_smem_bytes(p.n, restart, precision);
  const int64_t slot = gmres_slot_doubles(p.n, p.num_pos, restart, precision);
  if (precision == 0)
    hipLaunchKernelGGL(k_gmres, dim3((unsigned)grid), dim3(kGmresThreads), smem, st, p, nnz,
                       values, rhs, x0, B, restart, max_iter, tol, jacobi, x, iters, res, conv, ws,
                       slot);
  else
    hipLaunchKernelGGL(k_gmres_ir, dim3((unsigned)grid), dim3(kGmresThreads), smem, st, p, nnz,
                       values, rhs, x0, B, restart, max_iter, tol, jacobi, x, iters, res, conv, ws,
                       slot);
  return hipGetLastError();
}

}  // namespace spx
