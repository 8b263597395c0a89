// newton_kernel.h — the fused Newton DC-operating-point kernel template,
// instantiated per (system size bucket N, precision) in newton_*.hip.
//
// One 64-wide wavefront per circuit instance (one workgroup), so each
// instance's convergence mask is its own control flow — a converged
// instance's wave retires and the dispatcher backfills the CU with the next
// instance.  Per Newton iteration, all on-chip:
//   1. device evaluation: lanes over elements (diode / MOSFET companion
//      models in FP64, node voltages read from LDS);
//   2. assembly: lanes over the shared sparsity pattern, each summing its
//      entry's contributions in the reference's triplet order into the
//      pattern's CSR value array in LDS (the dense matrix is never stored:
//      a lane expands its row into VGPRs from the row's column bitmask);
//   3. linear solve: row-per-lane LU in VGPRs (FP64), or FP32 factors reused
//      across Newton iterations with FP64-residual iterative refinement
//      (residuals and the KCL test are sparse row products);
//   4. Δx convergence test (newton.rs:134-159) as a wave vote.
#pragma once

#include <hip/hip_runtime.h>

#include "assembly.h"
#include "device_common.h"
#include "lu_wave.h"
#include "program.h"

namespace spx {

// LDS per instance: pattern values + column indices, per-row column
// bitmasks, rhs, x, scratch x, element values, parameters, junction states.
inline size_t newton_smem_doubles(const DevProgram& g, int N) {
  return (size_t)g.num_pos + (g.num_pos + 1) / 2 + N + 3 * N + g.num_vals + g.P + g.num_elems;
}

// Exact: the FP64 LU compares every bit of the pivot candidates (the exact
// pass); otherwise the FP64 LU runs the fast comparison and hands an instance
// whose pivot a near-tie could change to the exact pass.
template <int N, bool IR, bool Exact = false>
struct Instance {
  const DevProgram& g;
  const DevNrCfg& cfg;
  int lane;
  double* J;      // num_pos pattern values, CSR order
  const int* jc;  // num_pos column indices
  uint32_t* rmask;  // 2 words per row: columns present in the row
  double* rhs;    // N
  double* xs;     // N current x
  double* xw;     // N scratch x (IR iterate / saved level solution)
  double* vals;   // num_vals
  double* pv;     // P parameters
  double* vst;    // num_elems pnjlim junction voltages
  int q0 = 0, q1 = 0;    // this lane's row: pattern entries [q0, q1)
  int nfact = 0;  // LU factorizations performed
  int nref = 0;   // refinement steps performed
  const TranCtx* tc = nullptr;  // transient context (null: DC sweep)
  // FP32+IR form: the FP32 factors, whether they are valid, and whether they
  // may be carried from one Newton solve to the next (transient steps: the
  // Jacobian moves little from step to step, and refinement against the
  // current one corrects the difference or asks for a refactorization)
  RowLU<N, float> lu32;
  bool have32 = false;
  bool reuse32 = false;

  __device__ __forceinline__ void set_reuse(bool v) { reuse32 = v; }

  // publishes the row masks (a barrier follows)
  __device__ __forceinline__ void init_row() {
    uint32_t m0 = 0, m1 = 0;
    if (lane < g.n) {
      q0 = g.row_ptr[lane];
      q1 = g.row_ptr[lane + 1];
      m0 = (uint32_t)g.rowmask[2 * lane];
      m1 = (uint32_t)g.rowmask[2 * lane + 1];
    }
    if (lane < N) {
      rmask[2 * lane] = m0;
      rmask[2 * lane + 1] = m1;
    }
  }

  // device evaluation + pattern assembly at xs; returns true if a pnjlim
  // limiter acted on any diode
  // from the lane-major tables: coalesced element loads, independent gather
  // ops (the same sums in the same order as the pos_off / contrib tables)
  __device__ __forceinline__ bool assemble(double gmin, double sf, int limiter) {
    bool lim = false;
    for (int t = 0; t < g.el_rounds; ++t) {
      const int* rec = g.el_lm + t * kElemLm * kWave + lane;
      if (rec[0] >= 0)
        eval_element_at<1, kWave>(rec, rec[kElemInts * kWave], xs, pv, vals, vst, sf, limiter,
                                  lim, tc);
    }
    if (lane == 0) vals[kValGmin] = gmin;
    __syncthreads();
    if (g.gath_k > 0) {
      gather_lanes(g, lane, J, rhs, vals);
    } else {
      for (int q = lane; q < g.num_pos; q += kWave)
        J[q] = gather_sum(g.pos_off, g.contrib, q, vals);
      for (int r = lane; r < g.n; r += kWave)
        rhs[r] = gather_sum(g.rhs_off, g.rhs_contrib, r, vals);
    }
    __syncthreads();
    return __any(lim);
  }

  // Expands this lane's row into registers; rows past n are identity rows.
  // J[q] at q == q1 is still inside the LDS block (the column indices follow
  // J), and is discarded when the column bit is clear.
  // The read of J[q] is kept unconditional (the empty asm pins it) so the
  // expansion is straight-line selects, not a branch per column.
  // Returns the row's chunk mask (RowLU::factor): the chunks of the pattern
  // columns, or of the identity entry.
  template <typename T>
  __device__ __forceinline__ uint32_t load_rows(T (&a)[N]) const {
    const bool pad = lane >= g.n;
    const int r = pad ? 0 : lane;  // rows past n (and lanes past N) take no pattern entry
    uint32_t mlo = pad ? 0u : rmask[2 * r], mhi = pad ? 0u : rmask[2 * r + 1];
    // opaque per call: otherwise the 64 column tests and the identity-row
    // constants are hoisted out of the Newton loop as 64 live lane masks and
    // 64 live values, which spill
    int diag = pad ? lane : -1;
    asm volatile("" : "+v"(mlo), "+v"(mhi), "+v"(diag));
    int q = q0;
    static_for<0, N>([&](auto JJ) __attribute__((always_inline)) {
      constexpr int j = JJ;
      const bool b = ((j < 32 ? mlo : mhi) >> (j & 31)) & 1u;
      double v = J[q];
      asm volatile("" : "+v"(v));
      a[j] = b ? (T)v : (j == diag ? T(1) : T(0));
      q += b ? 1 : 0;
      // at most 8 loads in flight: the scheduler would otherwise issue all N
      // first and hold 2N VGPRs of loaded values beside the row
      if constexpr ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    });
    static_assert(32 % kChunk == 0, "a chunk lies within one mask word");
    uint32_t cm = 0;
    static_for<0, (N + kChunk - 1) / kChunk>([&](auto C) __attribute__((always_inline)) {
      constexpr int c = C;
      const uint32_t w = (c * kChunk < 32 ? mlo : mhi) >> ((c * kChunk) & 31);
      if (w & ((1u << kChunk) - 1u)) cm |= 1u << c;
    });
    if (diag >= 0 && diag < N) cm = 1u << (diag / kChunk);
    return cm;
  }

  // J x for this lane's row (pattern order = column order)
  __device__ __forceinline__ double row_dot(const double* x, double r) const {
    for (int q = q0; q < q1; ++q) r = fma(-J[q], x[jc[q]], r);
    return r;
  }

  // x_new = J^{-1} rhs in FP64.  Returns 1 solved, 0 singular, 2 a near-tie
  // the exact pivot rule must decide (the instance goes to the exact pass).
  __device__ __forceinline__ int solve_f64(double& xnew) {
    ++nfact;
    RowLU<N, double> lu;
    const uint32_t cm = load_rows(lu.a);
    if (!lu.template factor<false, Exact>(lane, cm)) return (!Exact && lu.inexact) ? 2 : 0;
    xnew = lu.template solve<true>(lane < g.n ? rhs[lane] : 0.0, lane);
    return __all(lane >= g.n || isfinite(xnew)) ? 1 : 0;
  }

  // FP32 factors + FP64 iterative refinement, starting from the current x.
  // Stops on LAPACK dsgesv's backward-error test
  //   ||b - J x||_inf <= tol * ||J||_inf * ||x||_inf,  tol = ir_tol or sqrt(n)*2^-53,
  // i.e. once x is as good as an FP64 factorisation would give.  `fresh` says
  // the factors were computed from this J; returns 0 converged, 1 needs
  // refactoring (stale factors stopped contracting), 2 failed.
  __device__ __forceinline__ int refine(const RowLU<N, float>& lu, bool fresh, double& xnew) {
    const int n = g.n;
    const double b = lane < n ? rhs[lane] : 0.0;
    double x = lane < n ? xs[lane] : 0.0;
    double rowsum = 0.0;
    for (int q = q0; q < q1; ++q) rowsum += fabs(J[q]);
    const double anrm = wave_max(rowsum);
    const double tol = cfg.ir_tol > 0.0 ? cfg.ir_tol : sqrt((double)n) * 1.1102230246251565e-16;
    double prev = INFINITY;
    for (int s = 0; s <= cfg.ir_max_steps; ++s) {
      if (lane < n) xw[lane] = x;
      __syncthreads();
      const double r = row_dot(xw, b);
      __syncthreads();
      const double rn = wave_max(lane < n ? fabs(r) : 0.0);
      const double xn = wave_max(lane < n ? fabs(x) : 0.0);
      if (!(rn > tol * anrm * xn)) {  // converged (also catches r == 0)
        xnew = x;
        return __all(lane >= n || isfinite(x)) ? 0 : 2;
      }
      if (s == cfg.ir_max_steps) break;
      float d = lu.template solve<false>((float)r, lane);
      ++nref;
      x += (double)d;
      const double dn = wave_max(lane < n ? fabs((double)d) : 0.0);
      if (!(dn < 0.5 * prev) && s > 0) break;  // stopped contracting
      prev = dn;
    }
    return fresh ? 2 : 1;
  }

  // solve_newton_raphson at fixed gmin / source factor from xs; leaves the
  // final x in xs.  Returns 1 converged, 0 not converged, -1 singular.
  __device__ __forceinline__ int newton(double gmin, double sf, int& iters, int& fp64_escalations) {
    const int n = g.n, nn = g.num_nodes;
    for (int i = lane; i < g.num_elems; i += kWave) vst[i] = 0.0;
    __syncthreads();
    if (!reuse32) have32 = false;
    for (int it = 0; it < cfg.max_iterations; ++it) {
      const bool lim = assemble(gmin, sf, cfg.limiter);
      double xnew = 0.0;
      bool ok = true;
      if constexpr (IR) {
        int rc = 1;
        for (int attempt = 0; attempt < 2 && rc == 1; ++attempt) {
          bool fresh = !have32 || attempt > 0;
          if (fresh) {
            ++nfact;
            have32 = lu32.factor(lane, load_rows(lu32.a));
            if (!have32) { rc = 2; break; }
          }
          rc = refine(lu32, fresh, xnew);
        }
        if (rc != 0) {  // FP32 cannot resolve this system: hand the instance to FP64
          iters = it + 1;
          return -2;
        }
      } else {
        const int r = solve_f64(xnew);
        if (r == 2) {  // a near-tie pivot: the exact pass reruns the instance
          iters = it + 1;
          return -2;
        }
        ok = r == 1;
      }
      if (!ok) {
        iters = it + 1;
        return -1;
      }
      const double xo = lane < n ? xs[lane] : 0.0;
      bool pass = true;
      if (lane < n) {  // check_convergence, newton.rs:134-159
        double delta = fabs(xnew - xo);
        double tol = cfg.v_reltol * fmax(fabs(xnew), fabs(xo)) +
                     (lane < nn ? cfg.v_abstol : cfg.i_abstol);
        pass = !(delta > tol);
      }
      const bool conv = __all(pass) && !lim;
      __syncthreads();
      if (lane < n) xs[lane] = xnew;
      __syncthreads();
      if (conv) {
        iters = it + 1;
        return 1;
      }
    }
    iters = cfg.max_iterations;
    return 0;
  }

  __device__ __forceinline__ void save_x() {
    if (lane < g.n) xw[lane] = xs[lane];
    __syncthreads();
  }
  __device__ __forceinline__ void restore_x() {
    if (lane < g.n) xs[lane] = xw[lane];
    __syncthreads();
  }
  __device__ __forceinline__ void zero_x() {
    if (lane < g.n) xs[lane] = 0.0;
    __syncthreads();
  }

  // Linear circuit: one solve (sweep.rs:159-217), x left in xs (zeros if
  // singular).  Returns 0 solved, 1 singular, 2 FP32 could not resolve the
  // system (the FP64 pass takes the instance over).
  __device__ __forceinline__ int linear_solve() {
    assemble(0.0, 1.0, cfg.limiter);
    double xn = 0.0;
    int rc;
    if constexpr (IR) {
      int r = 1;
      for (int attempt = 0; attempt < 2 && r == 1; ++attempt) {
        const bool fresh = !(reuse32 && have32) || attempt > 0;
        if (fresh) {
          ++nfact;
          have32 = lu32.factor(lane, load_rows(lu32.a));
          if (!have32) {
            r = 2;
            break;
          }
        }
        r = refine(lu32, fresh, xn);
      }
      rc = r == 0 ? 0 : 2;
    } else {
      const int r = solve_f64(xn);
      rc = r == 1 ? 0 : (r == 2 ? 2 : 1);
    }
    if (lane < g.n) xs[lane] = rc == 0 ? xn : 0.0;
    __syncthreads();
    return rc;
  }

  __device__ __forceinline__ void write_result(const DevOut& out, int64_t inst, int status,
                                               int total, int flags, double res) const {
    const int n = g.n;
    if (lane < n) out.x[inst * n + lane] = status == SPX_SINGULAR ? 0.0 : xs[lane];
    if (lane == 0) {
      if (out.status) out.status[inst] = status;
      if (out.iterations) out.iterations[inst] = total;
      if (out.flags) out.flags[inst] = flags;
      if (out.residual) out.residual[inst] = res;
      if (out.factorizations) out.factorizations[inst] = nfact;
      if (out.refinements) out.refinements[inst] = nref;
    }
  }

  // Capacitor and inductor states from xs after a stage (cap_rule's and
  // ind_rule's rules; transient)
  __device__ __forceinline__ void cap_update(int rule, double h) {
    double* vp = const_cast<double*>(tc->vprev);
    double* ip = const_cast<double*>(tc->iprev);
    double* vq = const_cast<double*>(tc->vpp);
    for (int e = lane; e < g.num_elems; e += kWave) {
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
    for (int e = lane; e < g.num_elems; e += kWave) {
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
    lte = wave_max(l);
    ref = wave_max(r);
  }
  __device__ __forceinline__ void write_probes(const DevTran& tr, double* row) const {
    for (int p = lane; p < tr.num_probes; p += kWave) row[p] = xs[tr.probes[p]];
  }
  template <typename V>
  __device__ __forceinline__ void write_scalar(V* dst, V v) const {
    if (lane == 0) *dst = v;
  }

  __device__ __forceinline__ void write_handover(const DevOut& out, int64_t inst,
                                                 int flags) const {
    if (lane == 0) {
      out.status[inst] = SPX_ACTIVE;
      if (out.flags) out.flags[inst] = flags | SPX_FLAG_FP64_ESCALATED;
    }
  }

  // max_i |KCL_i| / tol_i at xs, devices unlimited (include/spicier_hip.h)
  __device__ __forceinline__ double kcl(double gmin) {
    assemble(gmin, 1.0, SPX_LIMIT_NONE);
    double q = 0.0;
    if (lane < g.n) {
      double r = -rhs[lane], mag = fabs(rhs[lane]);
      for (int q = q0; q < q1; ++q) {
        double t = J[q] * xs[jc[q]];
        r += t;
        mag = fmax(mag, fabs(t));
      }
      double tol = cfg.kcl_reltol * mag + (lane < g.num_nodes ? cfg.kcl_abstol : cfg.kcl_vntol);
      q = fabs(r) / tol;
      if (!(q == q)) q = INFINITY;
    }
    return wave_max(q);
  }
};

}  // namespace spx

#include "dcop_driver.h"

namespace spx {

// escalated_only: run only the instances a first pass handed over (status
// left SPX_ACTIVE), in FP64 with the exact pivot rule (Exact).
// Occupancy: the FP64 form at three waves per SIMD and the FP32+IR form at
// four, measured best on the 64-unknown sweep (two and five were slower).  The
// register budget this sets spills a few hundred bytes per wave outside the
// LU's inner steps; latency hiding across the extra waves is worth more.
//
// The exact-pivot pass (Exact) runs after a first pass that leaves the
// instances it hands over at SPX_ACTIVE: each wave reads the statuses of
// `scan` instances in one load and runs the active ones in turn, on a grid
// of B / scan waves.  After FP32+IR (escalated_only 1) hand-overs are rare
// and scan is 8: a grid of B one-wave workgroups that exit at once costs
// more to launch than the pass has work.  After the FP64 fast pass (2) the
// near-tie hand-overs are frequent enough that scan is 1, one instance per
// wave as before (8 serialised them, 4.85 -> 5.04 ms).
__host__ __device__ inline int exact_scan(int escalated_only) {
  return escalated_only == 1 ? 8 : 1;
}

template <int N, bool IR, bool Exact>
__global__ __launch_bounds__(64, IR ? 4 : 3) void k_newton_dense(DevProgram g, DevNrCfg cfg,
                                                     const double* __restrict__ params,
                                                     const double* __restrict__ x0, int64_t B,
                                                     DevOut out, int escalated_only) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x;
  double* J = smem;
  int* jc = reinterpret_cast<int*>(J + g.num_pos);
  uint32_t* rmask = reinterpret_cast<uint32_t*>(J + g.num_pos + (g.num_pos + 1) / 2);
  double* rhs = J + g.num_pos + (g.num_pos + 1) / 2 + N;
  double* xs = rhs + N;
  double* xw = xs + N;
  double* vals = xw + N;
  double* pv = vals + g.num_vals;
  double* vst = pv + g.P;
  const int n = g.n;
  auto run_one = [&](int64_t inst) {
    Instance<N, IR, Exact> I{g, cfg, lane, J, jc, rmask, rhs, xs, xw, vals, pv, vst};
    for (int q = lane; q < g.num_pos; q += kWave) jc[q] = g.pos_rc[q] >> 16;
    for (int p = lane; p < g.P; p += kWave) pv[p] = params[inst * g.P + p];
    if (lane == 0) { vals[kValOne] = 1.0; vals[kValGmin] = 0.0; }
    if (lane < N) xs[lane] = (x0 && lane < n) ? x0[inst * n + lane] : 0.0;
    __syncthreads();
    I.init_row();
    __syncthreads();
    run_dcop(I, g, cfg, x0 != nullptr, out, inst, escalated_only);
    if constexpr (Exact) __syncthreads();  // the next instance reuses the LDS
  };
  if constexpr (Exact) {
    const int scan = exact_scan(escalated_only);
    const int64_t base = (int64_t)blockIdx.x * scan;
    const int64_t mine = base + lane;
    const bool act = lane < scan && mine < B &&
                     (!escalated_only || out.status[mine] == SPX_ACTIVE);
    uint64_t m = (uint64_t)__ballot(act);
    while (m) {
      const int k = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      run_one(base + k);
    }
  } else {
    const int64_t inst = blockIdx.x;
    if (escalated_only && out.status[inst] != SPX_ACTIVE) return;
    run_one(inst);
  }
}

// Batched transient, one wave per instance: LDS as k_newton_dense plus the
// capacitor states (v_prev; with GENERAL also i_prev and v_prev_prev).  IR:
// FP32 factors carried across the time steps with FP64 refinement; an
// instance FP32 cannot resolve is left SPX_ACTIVE for the FP64 pass
// (escalated_only), which reruns it from t = 0.  GENERAL: the trapezoidal /
// TR-BDF2 / adaptive driver (run_transient<true>); otherwise Backward Euler.
template <int N, bool IR, bool GENERAL, bool Exact>
__global__ __launch_bounds__(64, IR ? 4 : 3) void k_transient_dense(DevProgram g, DevNrCfg cfg, DevTran tr,
                                                        const double* __restrict__ params,
                                                        int64_t B, DevOut out, DevTranOut tout,
                                                        int escalated_only) {
  extern __shared__ double smem[];
  const int64_t inst = blockIdx.x;
  const int lane = threadIdx.x;
  if (escalated_only && out.status[inst] != SPX_ACTIVE) return;
  double* J = smem;
  int* jc = reinterpret_cast<int*>(J + g.num_pos);
  uint32_t* rmask = reinterpret_cast<uint32_t*>(J + g.num_pos + (g.num_pos + 1) / 2);
  double* rhs = J + g.num_pos + (g.num_pos + 1) / 2 + N;
  double* xs = rhs + N;
  double* xw = xs + N;
  double* vals = xw + N;
  double* pv = vals + g.num_vals;
  double* vst = pv + g.P;
  double* vcap = vst + g.num_elems;
  double* icap = GENERAL ? vcap + g.num_elems : nullptr;
  double* vcap2 = GENERAL ? icap + g.num_elems : nullptr;
  Instance<N, IR, Exact> I{g, cfg, lane, J, jc, rmask, rhs, xs, xw, vals, pv, vst};
  TranCtx tc{0, 0.0, tr.h, vcap, icap, vcap2, 0, 0.0, 0.0, 0.0};
  I.tc = &tc;
  for (int q = lane; q < g.num_pos; q += kWave) jc[q] = g.pos_rc[q] >> 16;
  for (int p = lane; p < g.P; p += kWave) pv[p] = params[inst * g.P + p];
  for (int e = lane; e < (GENERAL ? 3 : 1) * g.num_elems; e += kWave) vcap[e] = 0.0;
  if (lane == 0) { vals[kValOne] = 1.0; vals[kValGmin] = 0.0; }
  if (lane < N) xs[lane] = 0.0;
  __syncthreads();
  I.init_row();
  __syncthreads();
  run_transient<GENERAL>(I, g, cfg, tr, tc, out, tout, inst);
}

// One dense transient kernel variant: op 1 sets its LDS attribute (not
// allowed while a stream is being captured into a graph), op 2 launches it.
template <int N, bool IR, bool GENERAL, bool Exact>
hipError_t transient_variant(int op, const DevProgram& g, const DevNrCfg& cfg, const DevTran& tr,
                             const double* params, int64_t B, const DevOut& out,
                             const DevTranOut& tout, hipStream_t st, int escalated_only) {
  const size_t sm =
      sizeof(double) * (newton_smem_doubles(g, N) + (GENERAL ? 3 : 1) * g.num_elems);
  if (op == 1)
    return hipFuncSetAttribute((const void*)k_transient_dense<N, IR, GENERAL, Exact>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  hipLaunchKernelGGL((k_transient_dense<N, IR, GENERAL, Exact>), dim3((unsigned)B), dim3(kWave),
                     sm, st, g, cfg, tr, params, B, out, tout, escalated_only);
  return hipGetLastError();
}

typedef hipError_t (*TranVariantFn)(int, const DevProgram&, const DevNrCfg&, const DevTran&,
                                    const double*, int64_t, const DevOut&, const DevTranOut&,
                                    hipStream_t, int);

// A variant as a plain function, so that each can be compiled in a file of
// its own (the 64-column register LU makes each variant slow to compile).
#define SPX_TRAN_VARIANT_FN(NAME, N, IR, GENERAL, EXACT)                                    \
  hipError_t NAME(int op, const DevProgram& g, const DevNrCfg& cfg, const DevTran& tr,      \
                  const double* params, int64_t B, const DevOut& out, const DevTranOut& tout, \
                  hipStream_t st, int esc) {                                                 \
    return transient_variant<N, IR, GENERAL, EXACT>(op, g, cfg, tr, params, B, out, tout, st, \
                                                    esc);                                    \
  }
#define SPX_TRAN_VARIANT_DECL(NAME)                                                         \
  hipError_t NAME(int, const DevProgram&, const DevNrCfg&, const DevTran&, const double*,   \
                  int64_t, const DevOut&, const DevTranOut&, hipStream_t, int);

// The transient of one size class in two passes over the same grid: the
// first kernel (FP32+IR, or FP64 with the fast pivot comparison) leaves every
// instance it hands over at SPX_ACTIVE, and the FP64 exact-pivot kernel
// reruns only those from t = 0 (every other workgroup exits on its first
// load); needs out.status.  v[general][0 FP64, 1 FP32+IR, 2 exact].
// phase bit 1: set the kernels' LDS attributes; bit 2: launch; bit 4: the
// FP32+IR form.  Backward-Euler fixed steps of circuits without inductors
// take the lean (general = 0) kernels.
inline hipError_t launch_transient_variants(const TranVariantFn (&v)[2][3], const DevProgram& g,
                                            const DevNrCfg& cfg, const DevTran& tr,
                                            const double* params, int64_t B, const DevOut& out,
                                            const DevTranOut& tout, hipStream_t st, int phase) {
  const int general = (tr.adaptive || tr.method != SPX_TRAN_BE || g.has_inductor) ? 1 : 0;
  const TranVariantFn first = v[general][(phase & 4) ? 1 : 0], exact = v[general][2];
  if (phase & 1) {
    hipError_t e = first(1, g, cfg, tr, params, B, out, tout, st, 0);
    if (e == hipSuccess) e = exact(1, g, cfg, tr, params, B, out, tout, st, 0);
    if (e != hipSuccess) return e;
  }
  if (!(phase & 2) || B == 0) return hipSuccess;
  if (!out.status) return hipErrorInvalidValue;
  hipError_t e = first(2, g, cfg, tr, params, B, out, tout, st, 0);
  if (e != hipSuccess) return e;
  return exact(2, g, cfg, tr, params, B, out, tout, st, 1);
}

template <int N, bool IR, bool Exact>
hipError_t launch_newton_impl(const DevProgram& g, const DevNrCfg& cfg, const double* params,
                              const double* x0, int64_t B, const DevOut& out,
                              int escalated_only, hipStream_t st) {
  size_t sm = sizeof(double) * newton_smem_doubles(g, N);
  hipError_t e = hipFuncSetAttribute((const void*)k_newton_dense<N, IR, Exact>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  const int scan = Exact ? exact_scan(escalated_only) : 1;
  const int64_t grid = (B + scan - 1) / scan;
  hipLaunchKernelGGL((k_newton_dense<N, IR, Exact>), dim3((unsigned)grid), dim3(kWave), sm, st, g,
                     cfg, params, x0, B, out, escalated_only);
  return hipGetLastError();
}

}  // namespace spx
