// kernels_sweep.hip — parameter sampling, the statistics reduction, and the
// dispatch of the Newton DC-operating-point kernel (newton_kernel.h) over its
// per-size / per-precision instantiations.
//
// Newton kernel: one 64-wide wavefront per circuit instance (one workgroup),
// so each instance's convergence mask is its own control flow — a converged
// instance's wave simply retires and the hardware dispatcher backfills the CU
// with the next instance.  Per Newton iteration, all on-chip:
//   1. device evaluation: lanes over elements (diode / MOSFET companion
//      models in FP64, reading node voltages from LDS);
//   2. assembly: lanes over the shared sparsity pattern, each summing its
//      entry's contributions in the reference's triplet order into a dense
//      LDS matrix (entries outside the pattern stay zero from the start);
//   3. linear solve: row-per-lane LU in VGPRs (FP64), or FP32 factors reused
//      across Newton iterations with FP64-residual iterative refinement;
//   4. Δx convergence test (newton.rs:134-159) as a wave vote.
// The stage logic (plain NR, Gmin stepping, source stepping) follows
// crates/spicier-solver/src/newton.rs:70-459; a circuit without nonlinear
// elements takes the single solve of solve_with_solver
// (crates/spicier-batched-sweep/src/sweep.rs:159-217).
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "lu_wave.h"
#include "program.h"

namespace spx {

// ------------------------------------------------------------- sampling
// One thread per (instance, slot) for the stateless generators; one thread
// per instance for the LCG, whose state is sequential over slots.
__global__ void k_sample_hash(const spx_param_spec* __restrict__ ps, int P, int gen, uint64_t seed,
                              int64_t B, int64_t offset, int nvar, double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * P) return;
  int64_t i = t / P;
  int p = (int)(t - i * P);
  const spx_param_spec q = ps[p];
  double v = q.a;
  if (q.kind != SPX_P_CONST) {
    if (gen == SPX_GEN_HASH) {
      uint32_t s = (uint32_t)(offset + i), r = (uint32_t)q.rng_index;
      switch (q.kind) {
        case SPX_P_GAUSS_REL:
          v = rclamp(q.a + hash_gaussian(seed, s, r) * q.b * q.a, q.c, q.d);
          break;
        case SPX_P_GAUSS:
          v = q.a + hash_gaussian(seed, s, r) * q.b;
          break;
        case SPX_P_UNIFORM:
          v = q.a + (q.b - q.a) * hash_uniform(seed, s, r);
          break;
        case SPX_P_LOGUNIFORM: {
          double la = log(q.a), lb = log(q.b);
          v = exp(la + (lb - la) * hash_uniform(seed, s, r));
          break;
        }
      }
    } else if (gen == SPX_GEN_CORNER) {
      // bit j of i selects min/max of the j-th varied slot (sweep.rs:204-227)
      int j = 0;
      for (int k = 0; k < p; ++k) j += ps[k].kind != SPX_P_CONST;
      v = ((i >> j) & 1) == 0 ? q.c : q.d;
    } else {  // SPX_GEN_LINEAR (sweep.rs:231-266)
      double tt = B > 1 ? (double)i / (double)(B - 1) : 0.5;
      v = q.c + tt * (q.d - q.c);
    }
  }
  out[i * P + p] = v;
}

__global__ void k_sample_lcg(const spx_param_spec* __restrict__ ps, int P, uint64_t seed,
                             int64_t B, int64_t offset, int nvar, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  // draws before instance i: two per varied slot per earlier instance
  Lcg g{lcg_skip(seed, (uint64_t)(2 * nvar) * (uint64_t)(offset + i))};
  for (int p = 0; p < P; ++p) {
    const spx_param_spec q = ps[p];
    double v = q.a;
    if (q.kind != SPX_P_CONST) v = rclamp(q.a + g.next_normal() * q.b * q.a, q.c, q.d);
    out[i * P + p] = v;
  }
}

// ------------------------------------------------------------- statistics
// Pass 1: block b reduces instances [b*chunk, (b+1)*chunk) for every node.
// Four waves take the rows in turn (wave w: rows w, w+4, ...), lane j the
// nodes j, j+64, ...: each step of a wave reads one contiguous row, and four
// rows' loads go out before their sums (the mask and status tests are
// wave-uniform scalar loads).  The waves' sums are then combined in wave
// order, so the result does not depend on timing.
constexpr int kStatWaves = 4;
__global__ __launch_bounds__(kStatWaves * kWave) void k_stats_partial(
    const double* __restrict__ v, const uint8_t* __restrict__ mask,
    const int* __restrict__ status, int64_t B, int n, int64_t chunk,
    double* __restrict__ part) {
  __shared__ double red[kStatWaves][5][kWave];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < B ? i0 + chunk : B;
  for (int j0 = 0; j0 < n; j0 += kWave) {
    const int j = j0 + lane;
    const bool col = j < n;
    double cnt = 0, sum = 0, sq = 0, mn = INFINITY, mx = -INFINITY;
    constexpr int U = 4;
    for (int64_t i = i0 + wave; i < i1; i += U * kStatWaves) {
      double x[U];
      bool take[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t r = i + (int64_t)u * kStatWaves;
        take[u] = r < i1 && (!mask || mask[r]) && (!status || status[r] == SPX_CONVERGED);
        x[u] = (take[u] && col) ? v[r * n + j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (take[u]) {
          cnt += 1;
          sum += x[u];
          sq += x[u] * x[u];
          mn = fmin(mn, x[u]);
          mx = fmax(mx, x[u]);
        }
    }
    red[wave][0][lane] = cnt;
    red[wave][1][lane] = sum;
    red[wave][2][lane] = sq;
    red[wave][3][lane] = mn;
    red[wave][4][lane] = mx;
    __syncthreads();
    if (wave == 0 && col) {
      for (int w = 1; w < kStatWaves; ++w) {
        cnt += red[w][0][lane];
        sum += red[w][1][lane];
        sq += red[w][2][lane];
        mn = fmin(mn, red[w][3][lane]);
        mx = fmax(mx, red[w][4][lane]);
      }
      double* o = part + ((int64_t)blockIdx.x * n + j) * 5;
      o[0] = cnt; o[1] = sum; o[2] = sq; o[3] = mn; o[4] = mx;
    }
    __syncthreads();
  }
}

// Pass 2: one workgroup per node merges the partials
// (StatisticsAccumulator::merge, statistics.rs:148-154): each thread folds a
// strided subset in order, then a fixed-shape tree combines the threads, so
// the result is deterministic run to run.
__global__ __launch_bounds__(256) void k_stats_merge(const double* __restrict__ part,
                                                     int nblocks, int n,
                                                     double* __restrict__ out) {
  __shared__ double s[5][256];
  const int j = blockIdx.x, t = threadIdx.x;
  double cnt = 0, sum = 0, sq = 0, mn = INFINITY, mx = -INFINITY;
  for (int b = t; b < nblocks; b += 256) {
    const double* p = part + ((int64_t)b * n + j) * 5;
    cnt += p[0]; sum += p[1]; sq += p[2];
    mn = fmin(mn, p[3]); mx = fmax(mx, p[4]);
  }
  s[0][t] = cnt; s[1][t] = sum; s[2][t] = sq; s[3][t] = mn; s[4][t] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s[0][t] += s[0][t + w];
      s[1][t] += s[1][t + w];
      s[2][t] += s[2][t + w];
      s[3][t] = fmin(s[3][t], s[3][t + w]);
      s[4][t] = fmax(s[4][t], s[4][t + w]);
    }
    __syncthreads();
  }
  if (t < 5) out[(int64_t)j * 5 + t] = s[t][0];
}

// ------------------------------------------------------------- launchers
#define SPX_DECLARE_NEWTON(N, IR, NAME)                                                   \
  hipError_t NAME(const DevProgram&, const DevNrCfg&, const double*, const double*, int64_t, \
                  const DevOut&, int, hipStream_t);
SPX_DECLARE_NEWTON(8, false, launch_newton_f64_8)
SPX_DECLARE_NEWTON(16, false, launch_newton_f64_16)
SPX_DECLARE_NEWTON(32, false, launch_newton_f64_32)
SPX_DECLARE_NEWTON(64, false, launch_newton_f64_64)
SPX_DECLARE_NEWTON(8, true, launch_newton_ir_8)
SPX_DECLARE_NEWTON(16, true, launch_newton_ir_16)
SPX_DECLARE_NEWTON(32, true, launch_newton_ir_32)
SPX_DECLARE_NEWTON(64, true, launch_newton_ir_64)
SPX_DECLARE_NEWTON(8, false, launch_newton_f64x_8)
SPX_DECLARE_NEWTON(16, false, launch_newton_f64x_16)
SPX_DECLARE_NEWTON(32, false, launch_newton_f64x_32)
SPX_DECLARE_NEWTON(64, false, launch_newton_f64x_64)

// Two passes over the same grid: the first (FP32+IR, or FP64 with the fast
// pivot comparison) leaves every instance it cannot resolve at SPX_ACTIVE,
// and the FP64 exact-pivot kernel solves only those (every other workgroup
// exits on its first load).  Needs out.status.
hipError_t launch_newton(const DevProgram& g, const DevNrCfg& cfg, const double* params,
                         const double* x0, int64_t B, const DevOut& out, hipStream_t st) {
  const bool ir = cfg.precision == SPX_PREC_F32_IR;
  typedef hipError_t (*Fn)(const DevProgram&, const DevNrCfg&, const double*, const double*,
                           int64_t, const DevOut&, int, hipStream_t);
  Fn f64 = g.n <= 8 ? launch_newton_f64_8 : g.n <= 16 ? launch_newton_f64_16
         : g.n <= 32 ? launch_newton_f64_32 : launch_newton_f64_64;
  Fn f32 = g.n <= 8 ? launch_newton_ir_8 : g.n <= 16 ? launch_newton_ir_16
         : g.n <= 32 ? launch_newton_ir_32 : launch_newton_ir_64;
  Fn exact = g.n <= 8 ? launch_newton_f64x_8 : g.n <= 16 ? launch_newton_f64x_16
           : g.n <= 32 ? launch_newton_f64x_32 : launch_newton_f64x_64;
  if (!out.status) return hipErrorInvalidValue;
  hipError_t e = (ir ? f32 : f64)(g, cfg, params, x0, B, out, 0, st);
  if (e != hipSuccess) return e;
  return exact(g, cfg, params, x0, B, out, ir ? 1 : 2, st);
}

hipError_t launch_sample(const spx_param_spec* d_specs, int P, int gen, uint64_t seed, int64_t B,
                         int64_t offset, int nvar, double* out, hipStream_t st) {
  if (B == 0 || P == 0) return hipSuccess;
  if (gen == SPX_GEN_LCG) {
    hipLaunchKernelGGL(k_sample_lcg, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, d_specs,
                       P, seed, B, offset, nvar, out);
  } else {
    int64_t total = B * P;
    hipLaunchKernelGGL(k_sample_hash, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       d_specs, P, gen, seed, B, offset, nvar, out);
  }
  return hipGetLastError();
}

int64_t stats_num_blocks(int64_t B) { return B < 1024 ? (B > 0 ? B : 1) : 1024; }

hipError_t launch_stats(const double* v, const uint8_t* mask, const int* status, int64_t B, int n,
                        double* part, double* out, hipStream_t st) {
  int64_t nb = stats_num_blocks(B);
  int64_t chunk = (B + nb - 1) / nb;
  if (chunk < 1) chunk = 1;
  hipLaunchKernelGGL(k_stats_partial, dim3((unsigned)nb), dim3(kStatWaves * kWave), 0, st, v, mask,
                     status, B, n, chunk, part);
  hipLaunchKernelGGL(k_stats_merge, dim3((unsigned)n), dim3(256), 0, st, part, (int)nb, n, out);
  return hipGetLastError();
}

}  // namespace spx
