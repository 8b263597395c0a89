// kernels_lu.hip — batched dense solve, the BatchedLuSolver::solve_batch path.
//
// n <= 64: one wavefront per system, row-per-lane LU in VGPRs (lu_wave.h).
//   The column-major input is read column by column, so the 64 lanes of a
//   wave read 64 consecutive doubles per load.
// n  > 64: one 256-thread workgroup per system on a global-memory copy of the
//   matrix (column k contiguous, so the pivot search and the multiplier
//   column are unit-stride); right-looking unblocked elimination.
// Both reproduce the reference CPU path's semantics
// (crates/spicier-batched-sweep/src/solver.rs:327-385 + faer_solver.rs:74-86):
// partial pivoting with the first maximum on ties, singular systems (zero
// pivot column or non-finite solution) get zeros and a flag.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "device_common.h"
#include "handle.h"
#include "lu_wave.h"

namespace spx {

template <int N>
__global__ __launch_bounds__(64) void k_solve_rows(const double* __restrict__ M,
                                                   const double* __restrict__ rhs, int n,
                                                   double* __restrict__ x,
                                                   uint8_t* __restrict__ sing) {
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const double* A = M + i * (int64_t)n * n;
  RowLU<N, double> lu;
  auto load = [&]() __attribute__((always_inline)) {
    static_for<0, N>([&](auto J) __attribute__((always_inline)) {
      constexpr int j = J;
      lu.a[j] = (j < n && lane < n) ? A[(int64_t)j * n + lane] : (j == lane ? 1.0 : 0.0);
    });
  };
  load();
  const double b = lane < n ? rhs[i * n + lane] : 0.0;
  // MNA matrices are sparse: the chunk masks of the loaded rows let the
  // factorization skip the zero chunks of each pivot row
  bool ok = lu.factor(lane, chunk_mask_of(lu.a));
  if (!ok && lu.inexact) {  // a near-tie: the exact pivot rule on the reloaded rows
    load();
    ok = lu.template factor<false, true>(lane, chunk_mask_of(lu.a));
  }
  double xv = 0.0;
  if (ok) {
    xv = lu.template solve<true>(b, lane);
    ok = __all(lane >= n || isfinite(xv));
  }
  if (lane < n) x[i * n + lane] = ok ? xv : 0.0;
  if (lane == 0 && sing) sing[i] = ok ? 0 : 1;
}

// getrf of one system (n <= N) in the register LU: factors in LAPACK's
// format (column-major, unit-lower L strictly below the diagonal, U on and
// above) and the 0-based row swapped with row k at step k.
template <int N>
__global__ __launch_bounds__(64) void k_factor_rows(const double* __restrict__ M, int n,
                                                    double* __restrict__ F,
                                                    int32_t* __restrict__ piv,
                                                    uint8_t* __restrict__ sing) {
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const double* A = M + i * (int64_t)n * n;
  RowLU<N, double> lu;
  auto load = [&]() __attribute__((always_inline)) {
    static_for<0, N>([&](auto J) __attribute__((always_inline)) {
      constexpr int j = J;
      lu.a[j] = (j < n && lane < n) ? A[(int64_t)j * n + lane] : (j == lane ? 1.0 : 0.0);
    });
  };
  load();
  bool ok = lu.template factor<true>(lane, chunk_mask_of(lu.a));
  if (!ok && lu.inexact) {
    load();
    ok = lu.template factor<true, true>(lane, chunk_mask_of(lu.a));
  }
  bool fin = true;
  static_for<0, N>([&](auto J) __attribute__((always_inline)) {
    constexpr int j = J;
    fin = fin && isfinite(lu.a[j]);
  });
  ok = ok && __all(lane >= n || fin);
  double* out = F + i * (int64_t)n * n;
  static_for<0, N>([&](auto J) __attribute__((always_inline)) {
    constexpr int j = J;
    if (j < n && lane < n) out[(int64_t)j * n + lane] = ok ? lu.a[j] : 0.0;
  });
  if (lane < n) piv[i * n + lane] = ok ? lu.ipiv : 0;
  if (lane == 0 && sing) sing[i] = ok ? 0 : 1;
}

// Right-looking unblocked partial-pivot LU of the column-major n x n matrix
// A in global memory by one 256-thread workgroup (column k contiguous, so the
// pivot search and the multiplier column are unit-stride): first maximum
// |a| on ties, physical row swaps (applied to y too when given), multipliers
// as a * (1/pivot), ipiv[k] the row swapped with row k when given.  Returns
// false at the first zero pivot column.  sv/si: 256 doubles / ints of LDS.
__device__ bool generic_factor(double* A, int n, double* y, int32_t* ipiv, double* sv, int* si,
                               int& s_p, double& s_pv) {
  const int tid = threadIdx.x;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int r = k + tid; r < n; r += 256) {
      double v = fabs(A[(int64_t)k * n + r]);
      if (v > best) { best = v; bi = r; }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) {
        double ov = sv[tid + s];
        int oi = si[tid + s];
        if (ov > sv[tid] || (ov == sv[tid] && oi < si[tid])) { sv[tid] = ov; si[tid] = oi; }
      }
      __syncthreads();
    }
    if (tid == 0) { s_p = si[0]; s_pv = sv[0]; }
    __syncthreads();
    const int p = s_p;
    if (!(s_pv > 0.0)) return false;
    if (tid == 0 && ipiv) ipiv[k] = p;
    if (p != k) {
      for (int c = tid; c < n; c += 256) {
        double t = A[(int64_t)c * n + k];
        A[(int64_t)c * n + k] = A[(int64_t)c * n + p];
        A[(int64_t)c * n + p] = t;
      }
      if (tid == 0 && y) { double t = y[k]; y[k] = y[p]; y[p] = t; }
    }
    __syncthreads();
    const double inv = 1.0 / A[(int64_t)k * n + k];
    const double yk = y ? y[k] : 0.0;
    for (int r = k + 1 + tid; r < n; r += 256) {
      double l = A[(int64_t)k * n + r] * inv;
      A[(int64_t)k * n + r] = l;
      if (y) y[r] = fma(-yk, l, y[r]);
    }
    __syncthreads();
    const int m = n - k - 1;
    const int64_t total = (int64_t)m * m;
    for (int64_t t = tid; t < total; t += 256) {
      int c = k + 1 + (int)(t / m), r = k + 1 + (int)(t % m);
      A[(int64_t)c * n + r] = fma(-A[(int64_t)c * n + k], A[(int64_t)k * n + r],
                                  A[(int64_t)c * n + r]);
    }
    __syncthreads();
  }
  return true;
}

__global__ __launch_bounds__(256) void k_solve_generic(const double* __restrict__ M,
                                                       const double* __restrict__ rhs, int n,
                                                       int64_t base, double* __restrict__ W,
                                                       double* __restrict__ Y,
                                                       double* __restrict__ x,
                                                       uint8_t* __restrict__ sing) {
  __shared__ double sv[256];
  __shared__ int si[256];
  __shared__ int s_p;
  __shared__ double s_pv;
  const int64_t li = blockIdx.x, gi = base + li;
  const int tid = threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  double* A = W + li * nn;
  double* y = Y + li * n;
  for (int64_t t = tid; t < nn; t += 256) A[t] = M[gi * nn + t];
  for (int t = tid; t < n; t += 256) y[t] = rhs[gi * n + t];
  __syncthreads();
  bool ok = generic_factor(A, n, y, nullptr, sv, si, s_p, s_pv);
  if (ok) {
    for (int k = n - 1; k >= 0; --k) {
      if (tid == 0) y[k] = y[k] / A[(int64_t)k * n + k];
      __syncthreads();
      const double yk = y[k];
      for (int r = tid; r < k; r += 256) y[r] = fma(-yk, A[(int64_t)k * n + r], y[r]);
      __syncthreads();
    }
    int bad = 0;
    for (int r = tid; r < n; r += 256) bad |= !isfinite(y[r]);
    ok = !__syncthreads_or(bad);
  }
  for (int r = tid; r < n; r += 256) x[gi * n + r] = ok ? y[r] : 0.0;
  if (tid == 0 && sing) sing[gi] = ok ? 0 : 1;
}

// getrf in place on the output F (copied from M): factors and 0-based
// pivots of LAPACK's format, zeros for a singular system.
__global__ __launch_bounds__(256) void k_factor_generic(const double* __restrict__ M, int n,
                                                        double* __restrict__ F,
                                                        int32_t* __restrict__ piv,
                                                        uint8_t* __restrict__ sing) {
  __shared__ double sv[256];
  __shared__ int si[256];
  __shared__ int s_p;
  __shared__ double s_pv;
  const int64_t i = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  double* A = F + i * nn;
  int32_t* ip = piv + i * n;
  for (int64_t t = tid; t < nn; t += 256) A[t] = M[i * nn + t];
  __syncthreads();
  bool ok = generic_factor(A, n, nullptr, ip, sv, si, s_p, s_pv);
  int bad = 0;
  if (ok)
    for (int64_t t = tid; t < nn; t += 256) bad |= !isfinite(A[t]);
  ok = !__syncthreads_or(bad || !ok);
  if (!ok) {
    for (int64_t t = tid; t < nn; t += 256) A[t] = 0.0;
    for (int t = tid; t < n; t += 256) ip[t] = 0;
  }
  if (tid == 0 && sing) sing[i] = ok ? 0 : 1;
}

template <int N>
static void launch_factor_rows(const double* M, int n, int64_t B, double* F, int32_t* piv,
                               uint8_t* sing, hipStream_t st) {
  hipLaunchKernelGGL(k_factor_rows<N>, dim3((unsigned)B), dim3(kWave), 0, st, M, n, F, piv, sing);
}

hipError_t launch_factor_batch(const double* M, int n, int64_t B, double* F, int32_t* piv,
                               uint8_t* sing, hipStream_t st) {
  if (B == 0) return hipSuccess;
  if (n <= 8) launch_factor_rows<8>(M, n, B, F, piv, sing, st);
  else if (n <= 16) launch_factor_rows<16>(M, n, B, F, piv, sing, st);
  else if (n <= 32) launch_factor_rows<32>(M, n, B, F, piv, sing, st);
  else if (n <= 64) launch_factor_rows<64>(M, n, B, F, piv, sing, st);
  else
    hipLaunchKernelGGL(k_factor_generic, dim3((unsigned)B), dim3(256), 0, st, M, n, F, piv, sing);
  return hipGetLastError();
}

template <int N>
static void launch_rows(const double* M, const double* rhs, int n, int64_t B, double* x,
                        uint8_t* sing, hipStream_t st) {
  hipLaunchKernelGGL(k_solve_rows<N>, dim3((unsigned)B), dim3(kWave), 0, st, M, rhs, n, x, sing);
}

// Bytes of workspace the generic path needs for a chunk of `chunk` systems.
size_t solve_generic_workspace(int n, int64_t chunk) {
  return (size_t)chunk * ((size_t)n * n + n) * sizeof(double);
}

hipError_t launch_solve_batch(const double* M, const double* rhs, int n, int64_t B, double* x,
                              uint8_t* sing, double* work, int64_t chunk, hipStream_t st) {
  if (B == 0) return hipSuccess;
  if (n <= 8) launch_rows<8>(M, rhs, n, B, x, sing, st);
  else if (n <= 16) launch_rows<16>(M, rhs, n, B, x, sing, st);
  else if (n <= 32) launch_rows<32>(M, rhs, n, B, x, sing, st);
  else if (n <= 64) launch_rows<64>(M, rhs, n, B, x, sing, st);
  else {
    for (int64_t base = 0; base < B; base += chunk) {
      int64_t cnt = B - base < chunk ? B - base : chunk;
      double* W = work;
      double* Y = work + (size_t)chunk * n * n;
      hipLaunchKernelGGL(k_solve_generic, dim3((unsigned)cnt), dim3(256), 0, st, M, rhs, n, base,
                         W, Y, x, sing);
    }
  }
  return hipGetLastError();
}

}  // namespace spx

// ---------------------------------------------------------------------------
// The no-allocation dense solve: the generic path's LU workspace lives in a
// file-scope device array instead of a handle workspace, so the call makes
// no device allocation.  Each GPU holds its own copy of the array (from the
// library's load until the process ends); host code reaches it only through
// the address hipGetSymbolAddress returns on that device, taken once per
// device when a handle is created there.

namespace spx {

namespace {
constexpr size_t kNoAllocScratchBytes = 64ull << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// With `device` current: records the device address of this device's
// scratch array (once per device).
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(SPX_ERR_BACKEND_INIT, "device " + std::to_string(device) + " out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return SPX_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SPX_ERR_BACKEND_INIT,
                std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = p;
  return SPX_OK;
}

}  // namespace spx

using namespace spx;

extern "C" int spx_solve_batch_f64_device_noalloc(spx_handle* h, const double* dM,
                                                  const double* drhs, int64_t n, int64_t B,
                                                  double* dx, uint8_t* dsing) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (B == 0) return SPX_OK;
  if (!dM || !drhs || !dx) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  if (n < 1 || n > 4096)
    return fail(SPX_ERR_INVALID_DIMENSION, "n must be 1..4096, got " + std::to_string(n));
  if (B > spx_max_batch_per_launch())
    return fail(SPX_ERR_BATCH_TOO_LARGE, "batch " + std::to_string(B) + " exceeds " +
                                             std::to_string(spx_max_batch_per_launch()));
  int64_t chunk = 1;
  if (n > 64) {  // one system's matrix and rhs must fit the scratch array
    const size_t per = solve_generic_workspace((int)n, 1);
    if (per > kNoAllocScratchBytes)
      return fail(SPX_ERR_INVALID_DIMENSION,
                  "the no-allocation solve holds one system of at most " +
                      std::to_string(kNoAllocScratchBytes) + " bytes (n <= 2895), got n = " +
                      std::to_string(n));
    chunk = std::min<int64_t>(B, (int64_t)(kNoAllocScratchBytes / per));
  }
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> sl(g_scratch_mu);
    ptr = (h->device >= 0 && h->device < kMaxDevices) ? g_scratch[h->device] : nullptr;
  }
  if (!ptr) return fail(SPX_ERR_BACKEND_INIT, "scratch address of the device is not known");
  SPX_HIP(launch_solve_batch(dM, drhs, (int)n, B, dx, dsing, static_cast<double*>(ptr), chunk,
                             h->stream),
          SPX_ERR_BACKEND);
  return SPX_OK;
}

