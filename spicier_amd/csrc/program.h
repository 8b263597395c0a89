// program.h — the device-resident circuit program shared by host and kernels.
//
// The host compiles an spx_circuit once per sweep into flat int tables:
//   elems   : per element {kind, u0, u1, u2, p0..p6, vid} (u = unknown index,
//             -1 for ground; vid = first value slot the element writes)
//   order   : element evaluation order, grouped by kind to keep a wave's
//             lanes on one branch of the device-model switch
//   pos_*   : the shared sparsity pattern — every (row, col) any stamp can
//             touch, each with its contributions (value slot, sign) in the
//             reference's triplet push order (mna.rs:62-71), so the sum is
//             formed in the same order as to_dense_matrix
//   row_ptr : CSR row starts into the (row, col)-sorted pattern
//   rhs_*   : the same for the right-hand side
//   el_lm   : the elements again, lane-major for one 64-lane wave: element
//             order[t*64 + l] in lane l of round t, field f at
//             [(t*kElemLm + f)*64 + l] (field 12: its element index; kind -1
//             pads), so each field load is one coalesced 256-byte access
//   gath    : the pattern and rhs sums as one fixed-length program per lane
//             for one 64-lane wave: lane l sums entries q = l, l+64, ... then
//             rhs rows r = l, l+64, ...; op k of lane l at [k*64 + l] is
//             slot | neg << 14 | last << 15 | dst << 16 (dst: pattern entry,
//             or 0x8000 | rhs row), -1 pads; gath_k = 0 when a table would
//             not fit those fields
// Value slots 0 and 1 hold the constant 1.0 (voltage-source couplings) and
// the current Gmin.  Per-iteration values live in LDS; assembly is one
// gather per pattern entry — no atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace spx {

constexpr int kElemInts = 12;
constexpr int kElemLm = 13;  // lane-major element fields: the 12 above + element index
constexpr int kValOne = 0;
constexpr int kValGmin = 1;

struct DevProgram {
  int n;           // unknowns
  int num_nodes;
  int P;           // parameter slots
  int num_elems;
  int num_vals;
  int num_pos;
  int has_nonlinear;
  int has_inductor;  // transients then carry each inductor's current history
  int max_row_nnz;
  const int* elems;
  const int* order;
  const int* pos_rc;   // row | col << 16, sorted by (row, col)
  const int* pos_off;  // num_pos + 1
  const int* row_ptr;  // n + 1: row r's pattern entries are [row_ptr[r], row_ptr[r+1])
  const int* contrib;  // vid << 1 | negate
  const int* rhs_off;  // n + 1
  const int* rhs_contrib;
  int el_rounds;       // ceil(num_elems / 64)
  const int* el_lm;    // el_rounds * kElemLm * 64
  int gath_k;          // ops per lane (0: use pos_off / contrib)
  const int* gath;     // gath_k * 64
  const int* rowmask;  // 2n words: columns present in row r (rows < 64)
};

// Newton / stepping configuration as the kernels read it.
struct DevNrCfg {
  double v_abstol, v_reltol, i_abstol;
  int max_iterations;
  int gmin_stepping;
  double gmin_initial, gmin_final, gmin_factor;
  int gmin_attempts;
  int source_stepping;
  double src_initial, src_step, src_min_step;
  int src_attempts;
  int precision, limiter, warm;
  int ir_max_steps;
  double ir_tol;
  double kcl_reltol, kcl_abstol, kcl_vntol;
};

// Transient context: mode 0 is the operating point the transient starts
// from (capacitors open, PULSE sources at V1, waveforms.rs:150-160); mode 1
// a stage solved at time t (PULSE sources at eval_pulse(t)) with the
// capacitor companion `comp` (companion.rs:41-130), each capacitor C
// becoming G with I from its negative to its positive node:
//   0 Backward Euler (stamp_be)      G = C/h, I = G*vprev
//   1 trapezoidal (stamp_trap)       G = 2C/h, I = G*vprev + iprev
//   2 BDF2 stage (stamp_trbdf2_bdf2) G = C/bden, I = G*(a1*vprev + a2*vpp)
// h is the companion's step (gamma*h for the TR-BDF2 trapezoidal stage).
// vprev / iprev / vpp[e] are capacitor element e's CapacitorState.  An
// inductor L takes the companion of InductorState (companion.rs:197-243):
//   0  G = h/L,      I = iprev
//   1  G = h/(2L),   I = iprev + G*vprev
//   2  G = bden/L,   I = a1*iprev + a2*vpp
// with I flowing from its first node to its second, and vprev / iprev /
// vpp[e] its v_prev, i_prev and i_prev_prev.
struct TranCtx {
  int mode;
  double t, h;
  const double* vprev;
  const double* iprev;
  const double* vpp;
  int comp;
  double a1, a2, bden;
};

// TRBDF2_GAMMA (transient/types.rs:11).
constexpr double kTrGamma = 2.0 - M_SQRT2;

// BDF2-stage coefficients of stamp_trbdf2_bdf2 (companion.rs:100-124) for
// a full step h: bden = b0 * h2.
__host__ __device__ inline void bdf2_coeffs(double h, double& a1, double& a2, double& bden) {
  const double h2 = (1.0 - kTrGamma) * h;
  const double h1 = kTrGamma * h;
  const double rho = h2 / h1;
  const double denom = 1.0 + 2.0 * rho;
  a1 = (1.0 + rho) * (1.0 + rho) / denom;
  a2 = -rho * rho / denom;
  const double b0 = (1.0 + rho) / denom;
  bden = b0 * h2;
}

// CapacitorState update of one capacitor with voltage v after a stage
// (companion.rs:62-98): rule -1 takes the operating point (vprev only; iprev
// and vpp start at 0), 0 Backward Euler, 1 trapezoidal, 2 the TR-BDF2 final
// update, 3 update_trbdf2_intermediate; h is the full step.
__device__ __forceinline__ void cap_rule(int rule, double C, double v, double h, double& vp,
                                         double& ip, double& vq) {
  if (rule < 0) {
    vp = v;
    return;
  }
  if (rule == 3) {
    const double h_gamma = kTrGamma * h;
    vq = vp;
    ip = 2.0 * C / h_gamma * (v - vq) - ip;
    vp = v;
    return;
  }
  if (rule == 0) {
    ip = C / h * (v - vp);
  } else if (rule == 1) {
    ip = 2.0 * C / h * (v - vp) - ip;
  } else {
    const double alpha = (1.0 - kTrGamma) / (kTrGamma * (2.0 - kTrGamma));
    ip = C / h * ((1.0 + alpha) * v - (1.0 + 2.0 * alpha) * vp + alpha * vq);
  }
  vq = vp;
  vp = v;
}

// estimate_lte of one capacitor (companion.rs:132-150) for a trapezoidal
// step of h ending at voltage v.
__device__ __forceinline__ double cap_lte(double C, double v, double h, double vp, double ip) {
  const double i_trap = 2.0 * C / h * (v - vp) - ip;
  const double i_be = C / h * (v - vp);
  return fabs(i_trap - i_be) / 3.0;
}

// InductorState update (companion.rs:219-262) of one inductor with branch
// voltage v after a stage, the same rules as cap_rule: vp = v_prev, ip =
// i_prev, iq = i_prev_prev.  Rule -1 takes the operating point (v_prev and
// the DC branch current ib, solver.rs:60-75).  The TR-BDF2 final update reads
// v_prev in both history terms, and the intermediate one leaves v_prev at the
// stage voltage (solver.rs:199-203), as the reference does.
__device__ __forceinline__ void ind_rule(int rule, double L, double v, double ib, double h,
                                         double& vp, double& ip, double& iq) {
  if (rule < 0) {
    vp = v;
    ip = ib;
    return;
  }
  if (rule == 3) {
    const double h_gamma = kTrGamma * h;
    iq = ip;
    ip += h_gamma / (2.0 * L) * (v + vp);
    vp = v;
    return;
  }
  if (rule == 0) {
    ip += h / L * v;
  } else if (rule == 1) {
    ip += h / (2.0 * L) * (v + vp);
  } else {
    const double alpha = (1.0 - kTrGamma) / (kTrGamma * (2.0 - kTrGamma));
    ip += h / L * ((1.0 + alpha) * v - (1.0 + 2.0 * alpha) * vp + alpha * vp);
  }
  iq = ip;
  vp = v;
}

// estimate_lte of one inductor (companion.rs:264-277) for a trapezoidal
// step of h ending at branch voltage v.
__device__ __forceinline__ double ind_lte(double L, double v, double h, double vp) {
  const double di_trap = h / (2.0 * L) * (v + vp);
  const double di_be = h / L * v;
  return fabs(di_trap - di_be) / 3.0;
}

// Transient run: fixed steps (h, steps, method) or adaptive trapezoidal
// stepping (tstop, h_init, h_min, h_max, reltol, abstol, max_points), and
// the probed unknowns.
struct DevTran {
  double h;
  int steps;
  int num_probes;
  const int* probes;
  int method;    // SPX_TRAN_*
  int adaptive;  // 0: fixed steps
  double tstop, h_init, h_min, h_max, reltol, abstol;
  int max_points;
};

struct DevTranOut {
  double* waves;   // [B][points][num_probes], points = steps + 1 or max_points
  int* fail_step;  // [B] first step that did not converge, -1 if none
  double* times;   // adaptive: [B][max_points]
  int* num_points; // adaptive: [B]
  int* total_steps;
  int* rejected_steps;
  double* min_step;
  double* max_step;
};

struct DevOut {
  double* x;
  int* status;
  int* iterations;
  int* flags;
  double* residual;
  int* factorizations;
  int* refinements;
};

}  // namespace spx
