// lanes.hip — Backward-Euler transient with one lane per instance: the 64
// lanes of a wave are 64 instances of one circuit running the shared-symbolic
// LU plan (sparse_plan.h) in lockstep.
//
// For circuits whose elimination tree is a chain (ladders, RC lines) the
// factorisation and both triangular solves are sequences of dependent single
// operations.  A workgroup per instance leaves all but one lane idle through
// each such chain; here every lane carries its own instance through the same
// sequence, plan indices are uniform across the wave (scalar loads), and
// the per-instance values are interleaved lane by lane in a global workspace
// (value k of lane l at base[k * 64 + l]) so each access is one coalesced
// 512-byte transaction.  Newton convergence, stepping and step failures are
// per lane (divergent control flow); nothing is shared between lanes, so
// the kernel needs no barrier.
#include <hip/hip_runtime.h>

#include "assembly.h"
#include "dcop_driver.h"
#include "program.h"
#include "sparse_plan.h"

namespace spx {

constexpr int kLanes = 64;

struct LaneInstance {
  static constexpr int ST = kLanes;
  const DevProgram& g;
  const DevSparse& sp;
  const DevNrCfg& cfg;
  const double* pv;  // every array below: this lane's element k at [k * ST]
  double* vals;
  double* vst;
  double* av;
  double* lu;
  double* rhs;
  double* xs;
  double* xw;
  double* y;
  int nfact = 0;
  int nref = 0;
  const TranCtx* tc = nullptr;

  __device__ __forceinline__ double& L(int k) const { return lu[k * ST]; }
  __device__ __forceinline__ double& Y(int k) const { return y[k * ST]; }

  __device__ __forceinline__ bool assemble(double gmin, double sf, int limiter) {
    bool lim = false;
    for (int i = 0; i < g.num_elems; ++i)
      eval_element<ST>(g, g.order[i], xs, pv, vals, vst, sf, limiter, lim, tc);
    vals[kValGmin * ST] = gmin;
    for (int q = 0; q < g.num_pos; ++q) {
      const double v = gather_sum<ST>(g.pos_off, g.contrib, q, vals);
      av[q * ST] = v;
      L(sp.pos_lu[q]) = v;
    }
    for (int f = 0; f < sp.num_fill; ++f) L(sp.fill[f]) = 0.0;
    for (int r = 0; r < g.n; ++r) rhs[r * ST] = gather_sum<ST>(g.rhs_off, g.rhs_contrib, r, vals);
    return lim;
  }

  // The plan in order (levels ascending, entries in order): the same
  // operations as the level-parallel kernels, so the same rounding.
  __device__ __forceinline__ bool factor_solve() {
    ++nfact;
    const int n = g.n;
    for (int Lv = 0; Lv < sp.levels; ++Lv) {
      for (int t = sp.pa_ptr[Lv]; t < sp.pa_ptr[Lv + 1]; ++t) {
        const int li = sp.pa_l[t];
        L(li) = L(li) * (1.0 / L(sp.pa_d[t]));
      }
      for (int t = sp.tb_ptr[Lv]; t < sp.tb_ptr[Lv + 1]; ++t) {
        const int ti = sp.tb_t[t];
        double v = L(ti);
        for (int c = sp.tb_c[t]; c < sp.tb_c[t + 1]; ++c) v = fma(-L(sp.cb_l[c]), L(sp.cb_u[c]), v);
        L(ti) = v;
      }
    }
    for (int k = 0; k < n; ++k)
      if (L(sp.diag[k]) == 0.0) return false;
    for (int k = 0; k < n; ++k) Y(k) = rhs[sp.perm[k] * ST];
    for (int r = 0; r < n; ++r) {  // forward, level order
      const int i = sp.lev_rows[r];
      const int u0 = sp.ut_ptr[i], nl = sp.ut_ptr[i + 1] - u0;
      const int l0 = sp.diag[i] - nl;
      double v = Y(i);
      for (int t = 0; t < nl; ++t) v = fma(-L(l0 + t), Y(sp.ut_row[u0 + t]), v);
      Y(i) = v;
    }
    for (int r = 0; r < n; ++r) {  // backward, level order from the root
      const int i = sp.blev_rows[r];
      const int s0 = sp.col_ptr[i], ns = sp.col_ptr[i + 1] - s0;
      const int d = sp.diag[i];
      double v = Y(i);
      for (int t = 0; t < ns; ++t) v = fma(-L(d + 1 + t), Y(sp.s_row[s0 + t]), v);
      Y(i) = v * (1.0 / L(d));
    }
    for (int k = 0; k < n; ++k)
      if (!isfinite(Y(k))) return false;
    return true;
  }

  __device__ __forceinline__ int newton(double gmin, double sf, int& iters, int& fp64_escalations) {
    const int n = g.n, nn = g.num_nodes;
    for (int i = 0; i < g.num_elems; ++i) vst[i * ST] = 0.0;
    for (int it = 0; it < cfg.max_iterations; ++it) {
      const bool lim = assemble(gmin, sf, cfg.limiter);
      if (!factor_solve()) {
        iters = it + 1;
        return -1;
      }
      bool pass = true;
      for (int k = 0; k < n; ++k) {  // check_convergence, newton.rs:134-159
        const int u = sp.perm[k];
        const double xn = Y(k), xo = xs[u * ST];
        const double tol = cfg.v_reltol * fmax(fabs(xn), fabs(xo)) +
                           (u < nn ? cfg.v_abstol : cfg.i_abstol);
        pass = pass && !(fabs(xn - xo) > tol);
      }
      for (int k = 0; k < n; ++k) xs[sp.perm[k] * ST] = Y(k);
      if (pass && !lim) {
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
    for (int k = 0; k < g.n; ++k) xs[sp.perm[k] * ST] = ok ? Y(k) : 0.0;
    return ok ? 0 : 1;
  }

  __device__ __forceinline__ void save_x() {
    for (int i = 0; i < g.n; ++i) xw[i * ST] = xs[i * ST];
  }
  __device__ __forceinline__ void restore_x() {
    for (int i = 0; i < g.n; ++i) xs[i * ST] = xw[i * ST];
  }
  __device__ __forceinline__ void zero_x() {
    for (int i = 0; i < g.n; ++i) xs[i * ST] = 0.0;
  }

  __device__ __forceinline__ double kcl(double gmin) {
    assemble(gmin, 1.0, SPX_LIMIT_NONE);
    double q = 0.0;
    for (int r = 0; r < g.n; ++r) {
      double res = -rhs[r * ST], mag = fabs(rhs[r * ST]);
      for (int e = g.row_ptr[r]; e < g.row_ptr[r + 1]; ++e) {
        const double t = av[e * ST] * xs[(g.pos_rc[e] >> 16) * ST];
        res += t;
        mag = fmax(mag, fabs(t));
      }
      const double tol = cfg.kcl_reltol * mag + (r < g.num_nodes ? cfg.kcl_abstol : cfg.kcl_vntol);
      double qr = fabs(res) / tol;
      if (!(qr == qr)) qr = INFINITY;
      q = fmax(q, qr);
    }
    return q;
  }

  __device__ __forceinline__ void set_reuse(bool) {}  // FP64 factors: nothing carried over
  // Capacitor and inductor states from xs after a stage (cap_rule's and
  // ind_rule's rules; transient)
  __device__ __forceinline__ void cap_update(int rule, double h) {
    double* vp = const_cast<double*>(tc->vprev);
    double* ip = const_cast<double*>(tc->iprev);
    double* vq = const_cast<double*>(tc->vpp);
    for (int e = 0; e < g.num_elems; ++e) {
      const int* el = g.elems + e * kElemInts;
      if (el[0] == SPX_EL_CAPACITOR) {
        const double v = (el[1] >= 0 ? xs[el[1] * ST] : 0.0) - (el[2] >= 0 ? xs[el[2] * ST] : 0.0);
        cap_rule(rule, pv[el[4] * ST], v, h, vp[e * ST], ip[e * ST], vq[e * ST]);
      } else if (el[0] == SPX_EL_INDUCTOR && ip) {  // el[3]: its branch row
        const double v = (el[1] >= 0 ? xs[el[1] * ST] : 0.0) - (el[2] >= 0 ? xs[el[2] * ST] : 0.0);
        ind_rule(rule, pv[el[4] * ST], v, xs[el[3] * ST], h, vp[e * ST], ip[e * ST], vq[e * ST]);
      }
    }
  }
  // Largest LTE of a trapezoidal step of h over the capacitors and inductors,
  // and the largest capacitor |v| and inductor |i_prev| (adaptive)
  __device__ __forceinline__ void cap_lte_max(double h, double& lte, double& ref) {
    double l = 0.0, r = 0.0;
    for (int e = 0; e < g.num_elems; ++e) {
      const int* el = g.elems + e * kElemInts;
      if (el[0] == SPX_EL_CAPACITOR) {
        const double v = (el[1] >= 0 ? xs[el[1] * ST] : 0.0) - (el[2] >= 0 ? xs[el[2] * ST] : 0.0);
        l = fmax(l, cap_lte(pv[el[4] * ST], v, h, tc->vprev[e * ST], tc->iprev[e * ST]));
        r = fmax(r, fabs(v));
      } else if (el[0] == SPX_EL_INDUCTOR) {  // the reference takes |i_prev| here
        const double v = (el[1] >= 0 ? xs[el[1] * ST] : 0.0) - (el[2] >= 0 ? xs[el[2] * ST] : 0.0);
        l = fmax(l, ind_lte(pv[el[4] * ST], v, h, tc->vprev[e * ST]));
        r = fmax(r, fabs(tc->iprev[e * ST]));
      }
    }
    lte = l;
    ref = r;
  }
  __device__ __forceinline__ void write_probes(const DevTran& tr, double* row) const {
    for (int p = 0; p < tr.num_probes; ++p) row[p] = xs[tr.probes[p] * ST];
  }
  template <typename V>
  __device__ __forceinline__ void write_scalar(V* dst, V v) const { *dst = v; }

  __device__ __forceinline__ void write_result(const DevOut& out, int64_t inst, int status,
                                               int total, int flags, double res) const {
    const int n = g.n;
    for (int i = 0; i < n; ++i) out.x[inst * n + i] = status == SPX_SINGULAR ? 0.0 : xs[i * ST];
    if (out.status) out.status[inst] = status;
    if (out.iterations) out.iterations[inst] = total;
    if (out.flags) out.flags[inst] = flags;
    if (out.residual) out.residual[inst] = res;
    if (out.factorizations) out.factorizations[inst] = nfact;
    if (out.refinements) out.refinements[inst] = nref;
  }
  __device__ __forceinline__ void write_handover(const DevOut& out, int64_t inst,
                                                 int flags) const {
    out.status[inst] = SPX_ACTIVE;
    if (out.flags) out.flags[inst] = flags | SPX_FLAG_FP64_ESCALATED;
  }
};

// Per-lane doubles of a slot (times kLanes for the slot).
inline int64_t lanes_instance_doubles(const DevProgram& g, const DevSparse& sp) {
  return (int64_t)g.P + 4 * g.n + g.num_vals + 4 * g.num_elems + g.num_pos + sp.nnz_lu;
}

template <bool GENERAL>
__global__ __launch_bounds__(kLanes) void k_transient_lanes(DevProgram g, DevSparse sp,
                                                            DevNrCfg cfg, DevTran tr,
                                                            const double* __restrict__ params,
                                                            int64_t B, DevOut out,
                                                            DevTranOut tout, double* ws,
                                                            int64_t slot) {
  const int lane = threadIdx.x;
  double* base = ws + blockIdx.x * slot + lane;
  double* pv = base;
  double* xs = pv + (int64_t)g.P * kLanes;
  double* rhs = xs + g.n * kLanes;
  double* xw = rhs + g.n * kLanes;
  double* y = xw + g.n * kLanes;
  double* vals = y + g.n * kLanes;
  double* vst = vals + (int64_t)g.num_vals * kLanes;
  double* vcap = vst + (int64_t)g.num_elems * kLanes;  // vprev, iprev, vpp
  double* av = vcap + 3 * (int64_t)g.num_elems * kLanes;
  double* lu = av + (int64_t)g.num_pos * kLanes;
  for (int64_t blk = blockIdx.x; blk * kLanes < B; blk += gridDim.x) {
    const int64_t inst = blk * kLanes + lane;
    if (inst >= B) continue;
    for (int p = 0; p < g.P; ++p) pv[p * kLanes] = params[inst * g.P + p];
    for (int i = 0; i < g.n; ++i) xs[i * kLanes] = 0.0;
    for (int e = 0; e < 3 * g.num_elems; ++e) vcap[e * kLanes] = 0.0;
    vals[kValOne * kLanes] = 1.0;
    vals[kValGmin * kLanes] = 0.0;
    TranCtx tc{0, 0.0, tr.h, vcap, vcap + (int64_t)g.num_elems * kLanes,
               vcap + 2 * (int64_t)g.num_elems * kLanes, 0, 0.0, 0.0, 0.0};
    LaneInstance I{g, sp, cfg, pv, vals, vst, av, lu, rhs, xs, xw, y};
    I.tc = &tc;
    run_transient<GENERAL>(I, g, cfg, tr, tc, out, tout, inst);
  }
}

hipError_t lanes_transient_shape(const DevProgram& g, const DevSparse& sp, int64_t B, int* grid,
                                 int64_t* slot) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
           &per_cu, (const void*)k_transient_lanes<false>, kLanes, 0)) != hipSuccess)
    return e;
  if (per_cu < 1) per_cu = 1;
  const int64_t blocks = (B + kLanes - 1) / kLanes;
  int64_t gsz = std::min<int64_t>((int64_t)cus * per_cu, blocks);
  *grid = (int)std::max<int64_t>(gsz, 1);
  *slot = ((lanes_instance_doubles(g, sp) * kLanes) + 31) & ~int64_t(31);
  return hipSuccess;
}

hipError_t launch_transient_lanes(const DevProgram& g, const DevSparse& sp, const DevNrCfg& cfg,
                                  const DevTran& tr, const double* params, int64_t B,
                                  const DevOut& out, const DevTranOut& tout, double* ws, int grid,
                                  int64_t slot, hipStream_t st) {
  if (B == 0) return hipSuccess;
  if (tr.adaptive || tr.method != SPX_TRAN_BE || g.has_inductor)
    hipLaunchKernelGGL(k_transient_lanes<true>, dim3((unsigned)grid), dim3(kLanes), 0, st, g, sp,
                       cfg, tr, params, B, out, tout, ws, slot);
  else
    hipLaunchKernelGGL(k_transient_lanes<false>, dim3((unsigned)grid), dim3(kLanes), 0, st, g,
                       sp, cfg, tr, params, B, out, tout, ws, slot);
  return hipGetLastError();
}

}  // namespace spx
