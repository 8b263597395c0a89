// postproc.hip — the per-instance bookkeeping and result reductions around a
// batched sweep, on the device: convergence tracking and active-set
// compaction (crates/spicier-batched-sweep/src/convergence.rs), histograms,
// yield and streaming statistics (statistics.rs), the FP32 packed matrix
// layout (batch_layout.rs), and a pipelined host-buffer batched solve
// (pipeline.rs).  All of these are HBM-bound integer / byte / FP64 work: the
// kernels read each input element once, coalesced, and count with integer
// atomics, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "device_common.h"
#include "handle.h"

namespace spx {

hipError_t launch_stats(const double* v, const uint8_t* mask, const int* status, int64_t B, int n,
                        double* part, double* out, hipStream_t st);
int64_t stats_num_blocks(int64_t B);
hipError_t launch_solve_batch(const double* M, const double* rhs, int n, int64_t B, double* x,
                              uint8_t* sing, double* work, int64_t chunk, hipStream_t st);
size_t solve_generic_workspace(int n, int64_t chunk);

namespace {

constexpr int kThreads = 256;

unsigned grid_for(int64_t items, int per_block = kThreads) {
  int64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;  // grid-stride loops cover the rest
  return (unsigned)g;
}

template <typename T>
__global__ void k_fill(T* p, int64_t count, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------ convergence tracker
// Rows of 64 instances are staged through LDS (coalesced loads, padded stride
// against bank conflicts); thread t then walks row t in column order, so the
// residual's sum of squares is accumulated in the reference's order
// (convergence.rs:339-345) and the result does not depend on the launch.
constexpr int kRowTile = 64;

template <bool kResidual>
__global__ __launch_bounds__(kRowTile) void k_tracker_check(
    int* __restrict__ status, int64_t B, int n, const double* __restrict__ a,
    const double* __restrict__ x, double abstol, double reltol, double tol,
    unsigned long long* __restrict__ newly) {
  __shared__ double sa[kRowTile][kRowTile + 1];
  __shared__ double sx[kRowTile][kRowTile + 1];
  const int t = threadIdx.x;
  for (int64_t i0 = (int64_t)blockIdx.x * kRowTile; i0 < B; i0 += (int64_t)gridDim.x * kRowTile) {
    const int rows = B - i0 < kRowTile ? (int)(B - i0) : kRowTile;
    const int64_t me = i0 + t;
    const bool active = t < rows && status[me] == SPX_ACTIVE;
    bool pass = true;
    double sq = 0.0;
    for (int c0 = 0; c0 < n; c0 += kRowTile) {
      const int cols = n - c0 < kRowTile ? n - c0 : kRowTile;
      __syncthreads();
      for (int r = 0; r < rows; ++r) {  // row r, columns c0 + t: coalesced
        if (t < cols) {
          sa[r][t] = a[(i0 + r) * n + c0 + t];
          if (!kResidual) sx[r][t] = x[(i0 + r) * n + c0 + t];
        }
      }
      __syncthreads();
      if (active) {
        for (int j = 0; j < cols; ++j) {
          if (kResidual) {
            const double r = sa[t][j];
            sq += r * r;
          } else {  // check_convergence, convergence.rs:291-303
            const double delta = fabs(sa[t][j]);
            const double tj = abstol + reltol * fabs(sx[t][j]);
            if (delta > tj) pass = false;
          }
        }
      }
    }
    const bool conv = active && (kResidual ? sqrt(sq) < tol : pass);
    if (conv) status[me] = SPX_CONVERGED;
    const unsigned long long m = __ballot(conv);
    if (newly && t == 0 && m) atomicAdd(newly, (unsigned long long)__popcll(m));
  }
}

// increment_all_active (convergence.rs:197-205 via increment_iteration :184-190)
__global__ void k_tracker_increment(int* __restrict__ status, int* __restrict__ iterations,
                                    int64_t B, int max_iterations) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < B;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (status[i] != SPX_ACTIVE) continue;
    const int it = iterations[i] + 1;
    iterations[i] = it;
    if (it >= max_iterations) status[i] = SPX_FAILED;
  }
}

// increment_iteration on listed instances (convergence.rs:184-190): the count
// grows whatever the status (repeated indices count repeatedly); a second
// pass fails the active ones at the limit.
__global__ void k_tracker_increment_idx(int* __restrict__ iterations, const int64_t* __restrict__ idx,
                                        int64_t count) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&iterations[idx[k]], 1);
}

__global__ void k_tracker_limit_idx(int* __restrict__ status, const int* __restrict__ iterations,
                                    const int64_t* __restrict__ idx, int64_t count,
                                    int max_iterations) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx[k];
    if (iterations[i] >= max_iterations) atomicCAS(&status[i], (int)SPX_ACTIVE, (int)SPX_FAILED);
  }
}

// mark_converged / mark_failed / mark_singular (convergence.rs:150-182): an
// instance changes only if it is active; the compare-and-swap makes a
// repeated index succeed once, as the sequential calls do.
__global__ void k_tracker_mark(int* __restrict__ status, const int64_t* __restrict__ idx,
                               int64_t count, int to, unsigned long long* __restrict__ marked) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (atomicCAS(&status[idx[k]], (int)SPX_ACTIVE, to) == SPX_ACTIVE) atomicAdd(marked, 1ull);
  }
}

// summary (convergence.rs:360-387): converged, failed (+singular), active,
// the converged instances' iteration sum and maximum.
__global__ void k_tracker_summary(const int* __restrict__ status,
                                  const int* __restrict__ iterations, int64_t B,
                                  unsigned long long* __restrict__ acc) {
  unsigned long long conv = 0, fail = 0, act = 0, isum = 0;
  unsigned int imax = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < B;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int s = status[i];
    if (s == SPX_CONVERGED) {
      ++conv;
      const unsigned it = iterations ? (unsigned)iterations[i] : 0u;
      isum += it;
      imax = it > imax ? it : imax;
    } else if (s == SPX_FAILED || s == SPX_SINGULAR) {
      ++fail;
    } else if (s == SPX_ACTIVE) {
      ++act;
    }
  }
  if (conv) atomicAdd(acc + 0, conv);
  if (fail) atomicAdd(acc + 1, fail);
  if (act) atomicAdd(acc + 2, act);
  if (isum) atomicAdd(acc + 3, isum);
  if (imax) atomicMax(acc + 4, (unsigned long long)imax);
}

// ------------------------------------------------------- ordered selection
// active_indices / converged_indices / failed_indices (convergence.rs:209-238)
// as a three-pass stream compaction with ascending output: per-tile counts,
// one exclusive scan of the tile counts, then each tile writes its matches at
// its offset in lane order.
constexpr int kSelTile = 4096;  // instances per tile (16 per thread)

__device__ __forceinline__ bool selected(int s, int which) {
  switch (which) {
    case SPX_SELECT_ACTIVE: return s == SPX_ACTIVE;
    case SPX_SELECT_CONVERGED: return s == SPX_CONVERGED;
    default: return s == SPX_FAILED || s == SPX_SINGULAR;
  }
}

__global__ __launch_bounds__(kThreads) void k_select_count(const int* __restrict__ status, int64_t B,
                                                           int which, int64_t* __restrict__ counts) {
  __shared__ int wsum[kThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kSelTile;
  int c = 0;
  for (int k = threadIdx.x; k < kSelTile; k += kThreads) {
    const int64_t i = base + k;
    c += (i < B && selected(status[i], which)) ? 1 : 0;
  }
  c = (int)wave_sum((double)c);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kThreads / kWave; ++w) s += wsum[w];
    counts[blockIdx.x] = s;
  }
}

// One workgroup: exclusive scan of the tile counts in place; total to *count.
__global__ __launch_bounds__(kThreads) void k_select_scan(int64_t* __restrict__ counts,
                                                          int64_t tiles,
                                                          int64_t* __restrict__ count) {
  __shared__ int64_t part[kThreads];
  const int t = threadIdx.x;
  const int64_t per = (tiles + kThreads - 1) / kThreads;
  const int64_t b0 = t * per, b1 = b0 + per < tiles ? b0 + per : tiles;
  int64_t s = 0;
  for (int64_t b = b0; b < b1; ++b) s += counts[b];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int k = 0; k < kThreads; ++k) {
      const int64_t v = part[k];
      part[k] = run;
      run += v;
    }
    if (count) *count = run;
  }
  __syncthreads();
  int64_t run = part[t];
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t v = counts[b];
    counts[b] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kThreads) void k_select_write(const int* __restrict__ status, int64_t B,
                                                           int which,
                                                           const int64_t* __restrict__ offsets,
                                                           int64_t* __restrict__ out) {
  __shared__ int wsum[kThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kSelTile;
  int64_t run = offsets[blockIdx.x];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  // chunks of kThreads consecutive instances, in order
  for (int k0 = 0; k0 < kSelTile; k0 += kThreads) {
    const int64_t i = base + k0 + threadIdx.x;
    const bool f = i < B && selected(status[i], which);
    const unsigned long long m = __ballot(f);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int q = 0; q < kThreads / kWave; ++q) {
      before += q < w ? wsum[q] : 0;
      total += wsum[q];
    }
    const int rank = before + __popcll(m & ((1ull << lane) - 1ull));
    if (f) out[run + rank] = i;
    run += total;
    __syncthreads();
  }
}

// compact_active (convergence.rs:432-441): out[k] = data[idx[k]]
__global__ void k_compact_rows(const double* __restrict__ data, int64_t item,
                               const int64_t* __restrict__ idx, int64_t count,
                               double* __restrict__ out) {
  const int64_t total = count * item;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / item, j = e - k * item;
    out[e] = data[idx[k] * item + j];
  }
}

// expand_active (convergence.rs:453-471): out[idx[k]] = compacted[k] (the
// caller's fill of the rest runs first)
__global__ void k_expand_rows(const double* __restrict__ comp, int64_t item,
                              const int64_t* __restrict__ idx, int64_t count,
                              double* __restrict__ out) {
  const int64_t total = count * item;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / item, j = e - k * item;
    out[idx[k] * item + j] = comp[e];
  }
}

// ----------------------------------------------------- histogram and yield
// Histogram::from_sweep's first pass (statistics.rs:279-285): min and max of
// one column, NaN skipped as f64::min / f64::max skip it.  One workgroup.
__global__ __launch_bounds__(kThreads) void k_column_range(const double* __restrict__ x, int64_t B,
                                                           int n, int col,
                                                           double* __restrict__ range) {
  __shared__ double smin[kThreads], smax[kThreads];
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t i = threadIdx.x; i < B; i += kThreads) {
    const double v = x[i * n + col];
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double lo = smin[0], hi = smax[0];
    if (fabs(hi - lo) < 2.220446049250313e-16) {  // :288-292
      lo -= 0.5;
      hi += 0.5;
    }
    range[0] = lo;
    range[1] = hi;
  }
}

// Histogram::new's binning (statistics.rs:253-258): floor((v - min) / width)
// saturated to an integer as Rust's `as isize` does (NaN -> 0), then clamped
// to the bins.  Counts gather in LDS per workgroup when the bins fit, then
// one 64-bit atomic per non-empty bin.
constexpr int kLdsBins = 4096;

__device__ __forceinline__ int bin_of(double v, double lo, double width, int bins) {
  const double t = floor((v - lo) / width);
  if (!(t == t)) return 0;
  if (t < 0.0) return 0;
  if (t > (double)(bins - 1)) return bins - 1;
  return (int)t;
}

__global__ __launch_bounds__(kThreads) void k_histogram(const double* __restrict__ x, int64_t B,
                                                        int n, int col, int bins,
                                                        const double* __restrict__ range,
                                                        unsigned long long* __restrict__ counts) {
  __shared__ unsigned int sb[kLdsBins];
  const double lo = range[0], hi = range[1];
  const double width = (hi - lo) / (double)bins;
  const bool lds = bins <= kLdsBins;
  if (lds)
    for (int b = threadIdx.x; b < bins; b += kThreads) sb[b] = 0u;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < B;
       i += (int64_t)gridDim.x * kThreads) {
    const int b = bin_of(x[i * n + col], lo, width, bins);
    if (lds) atomicAdd(&sb[b], 1u);
    else atomicAdd(&counts[b], 1ull);
  }
  __syncthreads();
  if (lds)
    for (int b = threadIdx.x; b < bins; b += kThreads)
      if (sb[b]) atomicAdd(&counts[b], (unsigned long long)sb[b]);
}

struct DevYieldSpec {
  int node;
  double lo, hi;
};

// YieldSpec::passes / YieldAnalysis::all_pass (statistics.rs:352-355,389-393):
// per spec and for all specs together, the count of passing instances.
__global__ __launch_bounds__(kThreads) void k_yield(const double* __restrict__ x, int64_t B, int n,
                                                    const DevYieldSpec* __restrict__ specs, int ns,
                                                    unsigned long long* __restrict__ counts) {
  // counts[0] = all pass, counts[1 + s] = spec s passes
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < B;
       i += (int64_t)gridDim.x * kThreads) {
    bool all = true;
    for (int s = 0; s < ns; ++s) {
      const double v = x[i * n + specs[s].node];
      const bool p = v >= specs[s].lo && v <= specs[s].hi;
      all = all && p;
      const unsigned long long m = __ballot(p);
      if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(&counts[1 + s], (unsigned long long)__popcll(m));
    }
    const unsigned long long m = __ballot(all);
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(&counts[0], (unsigned long long)__popcll(m));
  }
}

// StatisticsAccumulator::merge (statistics.rs:148-154) of a chunk's [n][5]
// moments into running accumulators.
__global__ void k_stats_fold(double* __restrict__ acc, const double* __restrict__ chunk, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double* a = acc + (int64_t)j * 5;
  const double* c = chunk + (int64_t)j * 5;
  a[0] += c[0];
  a[1] += c[1];
  a[2] += c[2];
  a[3] = fmin(a[3], c[3]);
  a[4] = fmax(a[4], c[4]);
}

// -------------------------------------------------------------- batch layout
// pack_matrices_f32 (batch_layout.rs:163-187): column-major FP64 systems to
// row-major FP32 with padded rows (padding = 0).  64 x 64 tiles transposed
// through LDS so both the column-major read and the row-major write are
// coalesced.
__global__ __launch_bounds__(256) void k_pack_f32(const double* __restrict__ in, int n, int64_t B,
                                                  int stride, float* __restrict__ out) {
  __shared__ float tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  const int ctiles = (stride + 63) / 64;
  const int r0 = (blockIdx.x / ctiles) * 64, c0 = (blockIdx.x % ctiles) * 64;
  for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
    const double* A = in + b * (int64_t)n * n;
    float* O = out + b * (int64_t)n * stride;
    for (int cy = ty; cy < 64; cy += 4) {  // column c0 + cy, rows r0 + tx
      const int c = c0 + cy, r = r0 + tx;
      tile[cy][tx] = (c < n && r < n) ? (float)A[(int64_t)c * n + r] : 0.0f;
    }
    __syncthreads();
    for (int ry = ty; ry < 64; ry += 4) {  // row r0 + ry, columns c0 + tx
      const int r = r0 + ry, c = c0 + tx;
      if (r < n && c < stride) O[(int64_t)r * stride + c] = tile[tx][ry];
    }
    __syncthreads();
  }
}

template <typename S, typename D>
__global__ void k_convert(const S* __restrict__ in, int64_t count, D* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (D)in[i];
}

int check_handle(spx_handle* h) { return h ? SPX_OK : fail(SPX_ERR_INVALID_ARGUMENT, "null handle"); }

}  // namespace
}  // namespace spx

using namespace spx;

extern "C" {

int spx_tracker_check_convergence_device(spx_handle* h, int32_t* d_status, int64_t B, int32_t n,
                                         const double* d_changes, const double* d_solutions,
                                         double abstol, double reltol,
                                         int64_t* d_newly_converged) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  auto* nw = reinterpret_cast<unsigned long long*>(d_newly_converged);
  if (nw) hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, h->stream, nw, 1, 0ull);
  if (B > 0) {
    if (!d_status || (n > 0 && (!d_changes || !d_solutions)))
      return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
    hipLaunchKernelGGL(k_tracker_check<false>, dim3(grid_for(B, kRowTile)), dim3(kRowTile), 0,
                       h->stream, d_status, B, n, d_changes, d_solutions, abstol, reltol, 0.0, nw);
  }
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_tracker_check_residual_device(spx_handle* h, int32_t* d_status, int64_t B, int32_t n,
                                      const double* d_residuals, double tolerance,
                                      int64_t* d_newly_converged) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  auto* nw = reinterpret_cast<unsigned long long*>(d_newly_converged);
  if (nw) hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, h->stream, nw, 1, 0ull);
  if (B > 0) {
    if (!d_status || (n > 0 && !d_residuals)) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
    hipLaunchKernelGGL(k_tracker_check<true>, dim3(grid_for(B, kRowTile)), dim3(kRowTile), 0,
                       h->stream, d_status, B, n, d_residuals, nullptr, 0.0, 0.0, tolerance, nw);
  }
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

// Copies host instance indices to the "pp_idx" workspace after checking them.
static int upload_indices(spx_handle* h, const int64_t* indices, int64_t count, int64_t B,
                          int64_t** d_idx) {
  for (int64_t k = 0; k < count; ++k)
    if (indices[k] < 0 || indices[k] >= B)
      return fail(SPX_ERR_INVALID_DIMENSION, "index " + std::to_string(indices[k]) +
                                                 " outside the batch of " + std::to_string(B));
  void* w = nullptr;
  if (int rc = workspace(h, "pp_idx", sizeof(int64_t) * std::max<int64_t>(count, 1), &w)) return rc;
  SPX_HIP(hipMemcpyAsync(w, indices, sizeof(int64_t) * count, hipMemcpyHostToDevice, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  *d_idx = static_cast<int64_t*>(w);
  return SPX_OK;
}

int spx_tracker_increment_device(spx_handle* h, int32_t* d_status, int32_t* d_iterations,
                                 int64_t B, int32_t max_iterations, const int64_t* indices,
                                 int64_t count) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || count < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (B == 0 || (indices && count == 0)) return SPX_OK;
  if (!d_status || !d_iterations) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  if (!indices) {
    hipLaunchKernelGGL(k_tracker_increment, dim3(grid_for(B)), dim3(kThreads), 0, h->stream,
                       d_status, d_iterations, B, max_iterations);
  } else {
    int64_t* di = nullptr;
    if (int rc = upload_indices(h, indices, count, B, &di)) return rc;
    hipLaunchKernelGGL(k_tracker_increment_idx, dim3(grid_for(count)), dim3(kThreads), 0,
                       h->stream, d_iterations, di, count);
    hipLaunchKernelGGL(k_tracker_limit_idx, dim3(grid_for(count)), dim3(kThreads), 0, h->stream,
                       d_status, d_iterations, di, count, max_iterations);
  }
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_tracker_mark_device(spx_handle* h, int32_t* d_status, int64_t B, const int64_t* indices,
                            int64_t count, int32_t new_status, int64_t* marked) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || count < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (new_status != SPX_CONVERGED && new_status != SPX_FAILED && new_status != SPX_SINGULAR)
    return fail(SPX_ERR_INVALID_ARGUMENT, "an instance can only be marked converged, failed or singular");
  if (marked) *marked = 0;
  if (count == 0) return SPX_OK;
  if (!d_status || !indices) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  int64_t* di = nullptr;
  if (int rc = upload_indices(h, indices, count, B, &di)) return rc;
  void* w = nullptr;
  if (int rc = workspace(h, "pp_marked", sizeof(unsigned long long), &w)) return rc;
  auto* m = static_cast<unsigned long long*>(w);
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, h->stream, m, 1, 0ull);
  hipLaunchKernelGGL(k_tracker_mark, dim3(grid_for(count)), dim3(kThreads), 0, h->stream, d_status,
                     di, count, (int)new_status, m);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  unsigned long long v = 0;
  SPX_HIP(hipMemcpyAsync(&v, m, sizeof(v), hipMemcpyDeviceToHost, h->stream), SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  if (marked) *marked = (int64_t)v;
  return SPX_OK;
}

int spx_tracker_summary_device(spx_handle* h, const int32_t* d_status, const int32_t* d_iterations,
                               int64_t B, spx_tracker_summary* out) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (!out) return fail(SPX_ERR_INVALID_ARGUMENT, "null summary");
  std::memset(out, 0, sizeof(*out));
  out->total_points = B;
  if (B == 0) return SPX_OK;
  if (!d_status) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void* acc = nullptr;
  if (int rc = workspace(h, "pp_summary", 5 * sizeof(unsigned long long), &acc)) return rc;
  auto* a = static_cast<unsigned long long*>(acc);
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, h->stream, a, 5, 0ull);
  hipLaunchKernelGGL(k_tracker_summary, dim3(grid_for(B)), dim3(kThreads), 0, h->stream, d_status,
                     d_iterations, B, a);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  unsigned long long v[5];
  SPX_HIP(hipMemcpyAsync(v, a, sizeof(v), hipMemcpyDeviceToHost, h->stream), SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  out->converged_count = (int64_t)v[0];
  out->failed_count = (int64_t)v[1];
  out->active_count = (int64_t)v[2];
  out->iteration_sum = (int64_t)v[3];
  out->max_iterations = (int32_t)v[4];
  out->average_iterations = v[0] ? (double)v[3] / (double)v[0] : 0.0;
  return SPX_OK;
}

int spx_select_indices_device(spx_handle* h, const int32_t* d_status, int64_t B, int32_t which,
                              int64_t* d_indices, int64_t* d_count) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (which != SPX_SELECT_ACTIVE && which != SPX_SELECT_CONVERGED && which != SPX_SELECT_FAILED)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown selection " + std::to_string(which));
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  if (B == 0) {
    if (d_count)
      hipLaunchKernelGGL(k_fill<int64_t>, dim3(1), dim3(64), 0, h->stream, d_count, 1, (int64_t)0);
    SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
    return SPX_OK;
  }
  if (!d_status || !d_indices) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  const int64_t tiles = (B + kSelTile - 1) / kSelTile;
  if (tiles > 0x7fffffff) return fail(SPX_ERR_BATCH_TOO_LARGE, "batch too large to select from");
  void* ws = nullptr;
  if (int rc = workspace(h, "pp_select", sizeof(int64_t) * tiles, &ws)) return rc;
  auto* offs = static_cast<int64_t*>(ws);
  hipLaunchKernelGGL(k_select_count, dim3((unsigned)tiles), dim3(kThreads), 0, h->stream, d_status,
                     B, which, offs);
  hipLaunchKernelGGL(k_select_scan, dim3(1), dim3(kThreads), 0, h->stream, offs, tiles, d_count);
  hipLaunchKernelGGL(k_select_write, dim3((unsigned)tiles), dim3(kThreads), 0, h->stream, d_status,
                     B, which, offs, d_indices);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_compact_rows_device(spx_handle* h, const double* d_data, int64_t item_size,
                            const int64_t* d_indices, int64_t count, double* d_out) {
  if (int rc = check_handle(h)) return rc;
  if (item_size < 0 || count < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (item_size == 0 || count == 0) return SPX_OK;
  if (!d_data || !d_indices || !d_out) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(count * item_size)), dim3(kThreads), 0,
                     h->stream, d_data, item_size, d_indices, count, d_out);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_expand_rows_device(spx_handle* h, const double* d_compacted, int64_t item_size,
                           const int64_t* d_indices, int64_t count, int64_t batch,
                           double fill_value, double* d_out) {
  if (int rc = check_handle(h)) return rc;
  if (item_size < 0 || count < 0 || batch < 0 || count > batch)
    return fail(SPX_ERR_INVALID_DIMENSION, "bad expand shape");
  if (item_size == 0 || batch == 0) return SPX_OK;
  if (!d_out || (count > 0 && (!d_compacted || !d_indices)))
    return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(batch * item_size)), dim3(kThreads), 0,
                     h->stream, d_out, batch * item_size, fill_value);
  if (count > 0)
    hipLaunchKernelGGL(k_expand_rows, dim3(grid_for(count * item_size)), dim3(kThreads), 0,
                       h->stream, d_compacted, item_size, d_indices, count, d_out);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_histogram_device(spx_handle* h, const double* d_solutions, int64_t B, int32_t n,
                         int32_t node, int32_t num_bins, double min, double max,
                         int32_t from_sweep, int64_t* d_counts, double* d_range) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || n < 1 || node < 0 || node >= n)
    return fail(SPX_ERR_INVALID_DIMENSION, "bad histogram shape");
  if (num_bins < 1) return fail(SPX_ERR_INVALID_ARGUMENT, "num_bins must be > 0");
  if (!from_sweep && !(max > min)) return fail(SPX_ERR_INVALID_ARGUMENT, "max must be > min");
  if (!d_counts || (B > 0 && !d_solutions)) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void* ws = nullptr;
  if (int rc = workspace(h, "pp_range", 2 * sizeof(double), &ws)) return rc;
  double* range = d_range ? d_range : static_cast<double*>(ws);
  auto* cnt = reinterpret_cast<unsigned long long*>(d_counts);
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(grid_for(num_bins)), dim3(kThreads), 0,
                     h->stream, cnt, (int64_t)num_bins, 0ull);
  if (from_sweep) {
    hipLaunchKernelGGL(k_column_range, dim3(1), dim3(kThreads), 0, h->stream, d_solutions, B,
                       (int)n, (int)node, range);
  } else {
    const double host_range[2] = {min, max};
    SPX_HIP(hipMemcpyAsync(range, host_range, sizeof(host_range), hipMemcpyHostToDevice,
                           h->stream),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  }
  if (B > 0)
    hipLaunchKernelGGL(k_histogram, dim3(grid_for(B)), dim3(kThreads), 0, h->stream, d_solutions,
                       B, (int)n, (int)node, (int)num_bins, range, cnt);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_yield_device(spx_handle* h, const double* d_solutions, int64_t B, int32_t n,
                     const spx_yield_spec* specs, int32_t num_specs, int64_t* d_counts) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || n < 1 || num_specs < 0) return fail(SPX_ERR_INVALID_DIMENSION, "bad yield shape");
  if (!d_counts || (num_specs > 0 && !specs) || (B > 0 && !d_solutions))
    return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::vector<DevYieldSpec> ds(num_specs);
  for (int s = 0; s < num_specs; ++s) {
    if (specs[s].node < 0 || specs[s].node >= n)
      return fail(SPX_ERR_INVALID_DIMENSION, "yield spec node out of range");
    ds[s] = DevYieldSpec{specs[s].node, specs[s].lower_limit, specs[s].upper_limit};
  }
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void* ws = nullptr;
  if (int rc = workspace(h, "pp_yield", sizeof(DevYieldSpec) * std::max(1, (int)num_specs), &ws))
    return rc;
  auto* cnt = reinterpret_cast<unsigned long long*>(d_counts);
  hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(kThreads), 0, h->stream, cnt,
                     (int64_t)num_specs + 1, 0ull);
  if (num_specs > 0) {
    SPX_HIP(hipMemcpyAsync(ws, ds.data(), sizeof(DevYieldSpec) * num_specs, hipMemcpyHostToDevice,
                           h->stream),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  }
  if (B > 0)
    hipLaunchKernelGGL(k_yield, dim3(grid_for(B)), dim3(kThreads), 0, h->stream, d_solutions, B,
                       (int)n, static_cast<const DevYieldSpec*>(ws), (int)num_specs, cnt);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_stats_accumulate_device(spx_handle* h, const double* d_values, const uint8_t* d_mask,
                                int64_t B, int32_t n, double* d_acc) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || n < 1) return fail(SPX_ERR_INVALID_DIMENSION, "bad statistics shape");
  if (!d_acc || (B > 0 && !d_values)) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  if (B == 0) return SPX_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void *part = nullptr, *chunk = nullptr;
  if (int rc = workspace(h, "stats_part2", sizeof(double) * 5 * n * stats_num_blocks(B), &part))
    return rc;
  if (int rc = workspace(h, "pp_chunk_stats", sizeof(double) * 5 * n, &chunk)) return rc;
  SPX_HIP(launch_stats(d_values, d_mask, nullptr, B, n, static_cast<double*>(part),
                       static_cast<double*>(chunk), h->stream),
          SPX_ERR_BACKEND);
  hipLaunchKernelGGL(k_stats_fold, dim3((n + 255) / 256), dim3(256), 0, h->stream, d_acc,
                     static_cast<const double*>(chunk), (int)n);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_stats_merge_device(spx_handle* h, double* d_acc, const double* d_other, int32_t n) {
  if (int rc = check_handle(h)) return rc;
  if (n < 1) return fail(SPX_ERR_INVALID_DIMENSION, "bad statistics shape");
  if (!d_acc || !d_other) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  hipLaunchKernelGGL(k_stats_fold, dim3((n + 255) / 256), dim3(256), 0, h->stream, d_acc, d_other,
                     (int)n);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_pack_matrices_f32_device(spx_handle* h, const double* d_matrices, int64_t n, int64_t B,
                                 int64_t row_stride, float* d_packed) {
  if (int rc = check_handle(h)) return rc;
  if (n < 0 || B < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (row_stride == 0) row_stride = (n + 31) / 32 * 32;  // BatchLayout::new, :43-46
  if (row_stride < n) return fail(SPX_ERR_INVALID_DIMENSION, "row stride below n");
  if (n > 0x7fffffff || row_stride > 0x7fffffff)
    return fail(SPX_ERR_INVALID_DIMENSION, "matrix too large");
  if (n == 0 || B == 0) return SPX_OK;
  if (!d_matrices || !d_packed) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  const int64_t tiles = ((n + 63) / 64) * ((row_stride + 63) / 64);
  if (tiles > 0x7fffffff) return fail(SPX_ERR_INVALID_DIMENSION, "matrix too large");
  const unsigned gy = (unsigned)std::min<int64_t>(B, 65535);
  hipLaunchKernelGGL(k_pack_f32, dim3((unsigned)tiles, gy), dim3(256), 0, h->stream, d_matrices,
                     (int)n, B, (int)row_stride, d_packed);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_pack_rhs_f32_device(spx_handle* h, const double* d_rhs, int64_t count, float* d_packed) {
  if (int rc = check_handle(h)) return rc;
  if (count < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (count == 0) return SPX_OK;
  if (!d_rhs || !d_packed) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(count)), dim3(kThreads), 0,
                     h->stream, d_rhs, count, d_packed);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_unpack_solutions_f64_device(spx_handle* h, const float* d_packed, int64_t count,
                                    double* d_solutions) {
  if (int rc = check_handle(h)) return rc;
  if (count < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (count == 0) return SPX_OK;
  if (!d_packed || !d_solutions) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(count)), dim3(kThreads), 0,
                     h->stream, d_packed, count, d_solutions);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

void spx_pipeline_config_auto(int64_t n, int64_t total_points, spx_pipeline_config* cfg) {
  // PipelineConfig::auto (pipeline.rs:68-87)
  if (!cfg) return;
  int64_t chunk = n < 50 ? 2000 : (n < 100 ? 1000 : 500);
  chunk = std::max<int64_t>(std::min<int64_t>(chunk, total_points / 4), 100);
  cfg->chunk_size = chunk;
  cfg->num_buffers = 2;
}

}  // extern "C"

namespace spx {
namespace {

// One slot of the pipelined solve: pinned staging for a chunk's inputs and
// outputs, its device buffers, a stream and a completion event.
struct PipeSlot {
  double *hM = nullptr, *hb = nullptr, *hx = nullptr;
  uint8_t* hs = nullptr;
  double *dM = nullptr, *db = nullptr, *dx = nullptr, *dw = nullptr;
  uint8_t* ds = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  int64_t start = -1, count = 0;  // chunk in flight (start < 0: none)
};

struct PipeSlots {
  std::vector<PipeSlot> s;
  ~PipeSlots() {
    for (auto& p : s) {
      if (p.st) (void)hipStreamSynchronize(p.st);
      (void)hipHostFree(p.hM);
      (void)hipHostFree(p.hb);
      (void)hipHostFree(p.hx);
      (void)hipHostFree(p.hs);
      (void)hipFree(p.dM);
      (void)hipFree(p.db);
      (void)hipFree(p.dx);
      (void)hipFree(p.dw);
      (void)hipFree(p.ds);
      if (p.ev) (void)hipEventDestroy(p.ev);
      if (p.st) (void)hipStreamDestroy(p.st);
    }
  }
};

}  // namespace
}  // namespace spx

extern "C" int spx_solve_batch_f64_pipelined(spx_handle* h, const double* M, const double* rhs,
                                             int64_t n, int64_t B, const spx_pipeline_config* cfg,
                                             double* x, uint8_t* sing, int64_t* num_singular,
                                             int64_t* chunks_processed) {
  if (int rc = check_handle(h)) return rc;
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (num_singular) *num_singular = 0;
  if (chunks_processed) *chunks_processed = 0;
  if (B == 0) return SPX_OK;
  if (!M || !rhs || !x) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  if (n < 1 || n > 4096)
    return fail(SPX_ERR_INVALID_DIMENSION, "n must be 1..4096, got " + std::to_string(n));
  spx_pipeline_config c;
  if (cfg) c = *cfg;
  else spx_pipeline_config_auto(n, B, &c);
  if (c.chunk_size < 1 || c.num_buffers < 1 || c.num_buffers > 8)
    return fail(SPX_ERR_INVALID_ARGUMENT, "chunk_size must be >= 1 and num_buffers 1..8");
  const int64_t chunk = std::min<int64_t>(std::min(c.chunk_size, B), spx_max_batch_per_launch());
  const int64_t nchunks = (B + chunk - 1) / chunk;
  const int nb = (int)std::min<int64_t>(c.num_buffers, nchunks);
  const size_t mat = sizeof(double) * (size_t)n * n;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  PipeSlots ps;
  ps.s.resize(nb);
  for (auto& p : ps.s) {
    SPX_HIP(hipHostMalloc((void**)&p.hM, mat * chunk, hipHostMallocDefault), SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipHostMalloc((void**)&p.hb, sizeof(double) * n * chunk, hipHostMallocDefault),
            SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipHostMalloc((void**)&p.hx, sizeof(double) * n * chunk, hipHostMallocDefault),
            SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipHostMalloc((void**)&p.hs, chunk, hipHostMallocDefault), SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipMalloc((void**)&p.dM, mat * chunk), SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipMalloc((void**)&p.db, sizeof(double) * n * chunk), SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipMalloc((void**)&p.dx, sizeof(double) * n * chunk), SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipMalloc((void**)&p.ds, chunk), SPX_ERR_MEMORY_ALLOC);
    if (n > 64)
      SPX_HIP(hipMalloc((void**)&p.dw, solve_generic_workspace((int)n, chunk)),
              SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipStreamCreateWithFlags(&p.st, hipStreamNonBlocking), SPX_ERR_BACKEND_INIT);
    SPX_HIP(hipEventCreateWithFlags(&p.ev, hipEventDisableTiming), SPX_ERR_BACKEND_INIT);
  }
  int64_t ns = 0;
  // drains slot p: waits for its chunk and copies the results out
  auto drain = [&](PipeSlot& p) -> int {
    if (p.start < 0) return SPX_OK;
    SPX_HIP(hipEventSynchronize(p.ev), SPX_ERR_TRANSFER);
    std::memcpy(x + p.start * n, p.hx, sizeof(double) * n * p.count);
    for (int64_t i = 0; i < p.count; ++i) ns += p.hs[i];
    if (sing) std::memcpy(sing + p.start, p.hs, p.count);
    p.start = -1;
    return SPX_OK;
  };
  // Chunk k goes to slot k mod nb.  While the GPU copies and solves chunk k,
  // the host stages chunk k+1 into the next slot's pinned buffer (after
  // draining that slot's previous chunk): host copies, both DMA directions
  // and the solve of different chunks overlap.
  for (int64_t k = 0; k < nchunks; ++k) {
    PipeSlot& p = ps.s[k % nb];
    if (int rc = drain(p)) return rc;
    const int64_t start = k * chunk, cnt = std::min(chunk, B - start);
    std::memcpy(p.hM, M + start * n * n, mat * cnt);
    std::memcpy(p.hb, rhs + start * n, sizeof(double) * n * cnt);
    SPX_HIP(hipMemcpyAsync(p.dM, p.hM, mat * cnt, hipMemcpyHostToDevice, p.st), SPX_ERR_TRANSFER);
    SPX_HIP(hipMemcpyAsync(p.db, p.hb, sizeof(double) * n * cnt, hipMemcpyHostToDevice, p.st),
            SPX_ERR_TRANSFER);
    SPX_HIP(launch_solve_batch(p.dM, p.db, (int)n, cnt, p.dx, p.ds, p.dw, cnt, p.st),
            SPX_ERR_BACKEND);
    SPX_HIP(hipMemcpyAsync(p.hx, p.dx, sizeof(double) * n * cnt, hipMemcpyDeviceToHost, p.st),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipMemcpyAsync(p.hs, p.ds, cnt, hipMemcpyDeviceToHost, p.st), SPX_ERR_TRANSFER);
    SPX_HIP(hipEventRecord(p.ev, p.st), SPX_ERR_BACKEND);
    p.start = start;
    p.count = cnt;
  }
  for (int64_t k = nchunks; k < nchunks + nb; ++k)
    if (int rc = drain(ps.s[k % nb])) return rc;
  if (num_singular) *num_singular = ns;
  if (chunks_processed) *chunks_processed = nchunks;
  return SPX_OK;
}
