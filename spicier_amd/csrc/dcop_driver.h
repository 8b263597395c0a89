// dcop_driver.h — the per-instance DC-operating-point state machine shared
// by the dense (register LU) and sparse (shared-symbolic LU) Newton kernels.
//
// An instance type provides: newton(gmin, source_factor, iters, escalations)
// -> 1 converged / 0 not / -1 singular / -2 hand over to FP64;
// linear_solve() -> 0 / 1 singular / 2 hand over; save_x / restore_x /
// zero_x; kcl(gmin) -> max_i |KCL_i| / tol_i; write_result; write_handover.
#pragma once

#include "program.h"

namespace spx {

// Outcome of one DC operating point: x is left in the instance.
struct DcopOutcome {
  int status, total, flags;
  double final_gmin;
  bool handover;  // FP32+IR could not resolve the system: the FP64 pass takes it
};

template <class Inst>
__device__ __forceinline__ DcopOutcome solve_dcop(Inst& I, const DevProgram& g,
                                                  const DevNrCfg& cfg, bool warm,
                                                  int escalated_only) {
  int total = 0, it = 0, esc = 0, status = SPX_FAILED;
  int flags = (warm ? SPX_FLAG_WARM_START : 0) |
              (escalated_only == 1 ? SPX_FLAG_FP64_ESCALATED : 0);
  double final_gmin = 0.0;
  bool handover = false;  // FP32+IR pass: leave this instance to the FP64 pass

  if (!g.has_nonlinear) {
    // linear circuit: one solve (sweep.rs:159-217)
    const int rc = I.linear_solve();
    handover = rc == 2;
    status = rc == 0 ? SPX_CONVERGED : SPX_SINGULAR;
    total = 1;
  } else {
    // Plain NR, then Gmin stepping (newton.rs:373-459), then source stepping
    // (newton.rs:261-356), as one state machine so the Newton loop has a
    // single, inlined call site.  Retries at one level restart from the same
    // solution and so replay the first attempt exactly; the first attempt
    // decides the level, and a level that does not converge counts its
    // iterations once per attempt, max_attempts_per_level times, as the
    // reference's total_iterations does (newton.rs:300-315, 411-426).  A
    // stage that fails returns its last converged level's solution (zeros if
    // none).
    enum { kPlain, kGmin, kSource, kDone };
    int stage = kPlain;
    double gm = 0.0, factor = 0.0, step = 0.0;
    auto enter_gmin = [&]() -> int {
      flags |= SPX_FLAG_GMIN_STEPPING;
      I.zero_x();
      I.save_x();
      gm = cfg.gmin_initial;
      if (gm >= cfg.gmin_final * 0.99) return kGmin;
      status = SPX_CONVERGED;  // no level to run: the stage returns converged zeros
      return kDone;
    };
    auto enter_source = [&]() -> int {
      flags |= SPX_FLAG_SOURCE_STEPPING;
      I.zero_x();
      I.save_x();
      factor = cfg.src_initial;
      step = cfg.src_step;
      if (factor <= 1.0 + 1e-9) return kSource;
      status = SPX_CONVERGED;
      final_gmin = 0.0;
      return kDone;
    };
    auto after_failure = [&]() -> int {
      if (stage == kPlain && cfg.gmin_stepping) return enter_gmin();
      if (stage != kSource && cfg.source_stepping) return enter_source();
      return kDone;
    };
    while (stage != kDone) {
      double lg = 0.0, lsf = 1.0;
      if (stage == kGmin) {
        lg = gm;
      } else if (stage == kSource) {
        lsf = fmin(factor, 1.0);
        I.restore_x();
      }
      const int st = I.newton(lg, lsf, it, esc);
      total += it;
      if (st == 0 && stage != kPlain)  // the level's replayed attempts
        total += (stage == kGmin ? cfg.gmin_attempts - 1 : cfg.src_attempts - 1) * it;
      if (st == -2) {
        handover = true;
        stage = kDone;
      } else if (stage == kPlain) {
        status = st == 1 ? SPX_CONVERGED : (st < 0 ? SPX_SINGULAR : SPX_FAILED);
        stage = status == SPX_CONVERGED ? kDone : after_failure();
      } else if (stage == kGmin) {
        if (st != 1) {
          I.restore_x();
          status = st < 0 ? SPX_SINGULAR : SPX_FAILED;
          final_gmin = 0.0;
          stage = after_failure();
        } else {
          I.save_x();
          final_gmin = gm;
          bool more = gm > cfg.gmin_final;
          if (more) {
            gm /= cfg.gmin_factor;
            if (gm < cfg.gmin_final) gm = cfg.gmin_final;
            more = gm >= cfg.gmin_final * 0.99;
          }
          if (!more) {
            I.restore_x();
            status = SPX_CONVERGED;
            stage = kDone;
          }
        }
      } else {  // kSource
        const double cur = fmin(factor, 1.0);
        bool finished = false;
        if (st != 1) {
          if (step > cfg.src_min_step) {
            step /= 2.0;
            factor -= step;
            finished = !(factor <= 1.0 + 1e-9);
          } else {
            I.restore_x();
            status = st < 0 ? SPX_SINGULAR : SPX_FAILED;
            stage = kDone;
          }
        } else {
          I.save_x();
          if (fabs(cur - 1.0) < 1e-9) {
            finished = true;
          } else {
            factor += step;
            if (step < cfg.src_step) step = fmin(step * 1.5, cfg.src_step);
            finished = !(factor <= 1.0 + 1e-9);
          }
        }
        if (finished) {
          I.restore_x();
          status = SPX_CONVERGED;
          final_gmin = 0.0;
          stage = kDone;
        }
      }
    }
  }
  return DcopOutcome{status, total, flags, final_gmin, handover};
}

template <class Inst>
__device__ __forceinline__ void run_dcop(Inst& I, const DevProgram& g, const DevNrCfg& cfg,
                                         bool warm, const DevOut& out, int64_t inst,
                                         int escalated_only) {
  DcopOutcome o = solve_dcop(I, g, cfg, warm, escalated_only);
  if (o.handover) {
    I.write_handover(out, inst, o.flags);
    return;
  }
  const double res = I.kcl(o.final_gmin);
  if (o.status == SPX_CONVERGED && !(res <= 1.0)) o.flags |= SPX_FLAG_KCL_FAIL;
  I.write_result(out, inst, o.status, o.total, o.flags, res);
}

// One stage of a transient step with the companion set in tc: linear
// circuits one solve (the reference's), nonlinear ones plain Newton from
// the current solution.  Returns 1 solved, -2 hand-over (FP32 factors
// cannot resolve the instance), anything else a failure (< 0 singular).
template <class Inst>
__device__ __forceinline__ int tran_stage(Inst& I, const DevProgram& g, int& it) {
  if (!g.has_nonlinear) {
    const int r = I.linear_solve();
    it = 1;
    return r == 0 ? 1 : (r == 2 ? -2 : -1);
  }
  int esc = 0;
  it = 0;
  return I.newton(0.0, 1.0, it, esc);
}

// Transient of one instance (transient/solver.rs:41-249, 461-676): the
// operating point (capacitors open, PULSE sources at V1) gives every
// capacitor v_prev (i_prev = 0), then either
//  * `steps` fixed steps of h solved at t = step * h with the method's
//    companion: Backward Euler, trapezoidal, or TR-BDF2 (a trapezoidal
//    stage of gamma*h then the BDF2 stage, both with the sources at t, the
//    capacitor states updated after each stage as the reference does); or
//  * adaptive trapezoidal steps with the sources at the start of the step,
//    each instance on its own time grid: a step whose largest capacitor LTE
//    exceeds max(abstol, reltol * max|v_cap|) while h > h_min is rejected
//    and retried with h scaled by clamp(sqrt(tol/lte), 0.1, 0.5); an
//    accepted step with LTE < tol/2 lets h grow by up to 1.5x; h stays in
//    [h_min, h_max] and is cut to end exactly at tstop.
// The reference's transient is linear (one solve per stage); a circuit
// with diodes or MOSFETs runs plain Newton at each stage from the latest
// solution.  Probes (and, adaptive, times) record t = 0 and every accepted
// point; the result row carries the final solution, the total Newton
// iterations (operating point included) and the FP64 KCL residual of the
// last stage.  An instance whose operating point or a stage does not
// converge, or that needs more than max_points points, stops with status
// FAILED (or SINGULAR) and fail_step set.
//
// Inductors (InductorState, companion.rs:158-302) start from their DC
// branch current and voltage and follow the same stage order with their own
// companions and updates; in adaptive runs their LTE joins the capacitors'
// and their |i_prev| the reference magnitude.
//
// GENERAL = false compiles the Backward-Euler fixed-step run of a circuit
// without inductors alone (capacitor state v_prev only): the kernels that
// serve it keep the register budget and occupancy of a lean step loop;
// GENERAL = true adds inductors, the other methods and adaptive stepping.
template <bool GENERAL, class Inst>
__device__ __forceinline__ void run_transient(Inst& I, const DevProgram& g, const DevNrCfg& cfg,
                                              const DevTran& tr, TranCtx& tc,
                                              const DevOut& out, const DevTranOut& tout,
                                              int64_t inst) {
  tc.mode = 0;
  tc.t = 0.0;
  DcopOutcome o = solve_dcop(I, g, cfg, false, 0);
  if (o.handover) {  // FP32 factors cannot resolve this instance: the FP64 pass reruns it
    I.write_handover(out, inst, o.flags);
    return;
  }
  int status = o.status, total = o.total, flags = o.flags, fail = -1;
  int points = 0, tried = 0, rejected = 0;
  double res = 0.0, hmin_used = INFINITY, hmax_used = 0.0;
  const int64_t stride = tr.adaptive ? tr.max_points : tr.steps + 1;
  const int64_t row0 = inst * stride;
  if (status == SPX_CONVERGED) {
    I.cap_update(-1, 0.0);
    if (tout.waves) I.write_probes(tr, tout.waves + row0 * tr.num_probes);
    if (tout.times) I.write_scalar(tout.times + row0, 0.0);
    points = 1;
    tc.mode = 1;
    I.set_reuse(true);
    if constexpr (!GENERAL) {
      const double h = tr.h;
      tc.comp = 0;
      tc.h = h;
      for (int step = 1; step <= tr.steps; ++step) {
        tc.t = step * h;
        int it = 0;
        const int st = tran_stage(I, g, it);
        total += it;
        if (st == -2) {
          I.write_handover(out, inst, flags);
          return;
        }
        if (st != 1) {
          status = st < 0 ? SPX_SINGULAR : SPX_FAILED;
          fail = step;
          break;
        }
        if (step == tr.steps) res = I.kcl(0.0);  // before the capacitors move on
        I.cap_update(-1, h);                     // Backward Euler reads v_prev only
        if (tout.waves) I.write_probes(tr, tout.waves + (row0 + step) * tr.num_probes);
        points = step + 1;
      }
      if (tr.steps == 0) res = I.kcl(o.final_gmin);
    } else if (!tr.adaptive) {
      const double h = tr.h;
      for (int step = 1; step <= tr.steps; ++step) {
        tc.t = step * h;
        int it = 0, st;
        if (tr.method == SPX_TRAN_TRBDF2) {
          tc.comp = 1;
          tc.h = kTrGamma * h;
          st = tran_stage(I, g, it);
          total += it;
          if (st == 1) {
            I.cap_update(3, h);
            tc.comp = 2;
            tc.h = h;
            bdf2_coeffs(h, tc.a1, tc.a2, tc.bden);
            st = tran_stage(I, g, it);
            total += it;
          }
          if (st == 1) {
            if (step == tr.steps) res = I.kcl(0.0);  // before the capacitors move on
            I.cap_update(2, h);
          }
        } else {
          tc.comp = tr.method == SPX_TRAN_TRAP ? 1 : 0;
          tc.h = h;
          st = tran_stage(I, g, it);
          total += it;
          if (st == 1) {
            if (step == tr.steps) res = I.kcl(0.0);
            I.cap_update(tc.comp, h);
          }
        }
        if (st == -2) {
          I.write_handover(out, inst, flags);
          return;
        }
        if (st != 1) {
          status = st < 0 ? SPX_SINGULAR : SPX_FAILED;
          fail = step;
          break;
        }
        if (tout.waves) I.write_probes(tr, tout.waves + (row0 + step) * tr.num_probes);
        points = step + 1;
      }
      if (tr.steps == 0) res = I.kcl(o.final_gmin);
    } else {
      double t = 0.0, h = tr.h_init;
      tc.comp = 1;
      while (t < tr.tstop) {
        h = h < tr.h_min ? tr.h_min : (h > tr.h_max ? tr.h_max : h);
        if (t + h > tr.tstop) h = tr.tstop - t;
        tc.t = t;  // the reference stamps the sources at the start of the step
        tc.h = h;
        int it = 0;
        const int st = tran_stage(I, g, it);
        total += it;
        ++tried;
        if (st == -2) {
          I.write_handover(out, inst, flags);
          return;
        }
        if (st != 1) {
          status = st < 0 ? SPX_SINGULAR : SPX_FAILED;
          fail = points;
          break;
        }
        double lte, ref;
        I.cap_lte_max(h, lte, ref);
        const double tol = fmax(tr.abstol, tr.reltol * ref);
        if (lte > tol && h > tr.h_min) {
          ++rejected;
          h *= fmax(fmin(sqrt(tol / lte), 0.5), 0.1);
          continue;
        }
        if (points >= tr.max_points) {
          status = SPX_FAILED;
          fail = points;
          break;
        }
        t += h;
        if (!(t < tr.tstop)) res = I.kcl(0.0);
        I.cap_update(1, h);
        hmin_used = fmin(hmin_used, h);
        hmax_used = fmax(hmax_used, h);
        if (tout.waves) I.write_probes(tr, tout.waves + (row0 + points) * tr.num_probes);
        if (tout.times) I.write_scalar(tout.times + row0 + points, t);
        ++points;
        if (lte < tol * 0.5 && h < tr.h_max)
          h *= fmin(fmin(sqrt(tol / fmax(lte, 1e-20)), 2.0), 1.5);
      }
    }
  } else {
    fail = 0;
    res = I.kcl(o.final_gmin);
  }
  if (status == SPX_CONVERGED && !(res <= 1.0)) flags |= SPX_FLAG_KCL_FAIL;
  I.write_result(out, inst, status, total, flags, res);
  if (tout.fail_step) I.write_scalar(tout.fail_step + inst, fail);
  if (tout.num_points) I.write_scalar(tout.num_points + inst, points);
  if (tout.total_steps) I.write_scalar(tout.total_steps + inst, tried);
  if (tout.rejected_steps) I.write_scalar(tout.rejected_steps + inst, rejected);
  if (tout.min_step) I.write_scalar(tout.min_step + inst, hmin_used);
  if (tout.max_step) I.write_scalar(tout.max_step + inst, hmax_used);
}

}  // namespace spx
