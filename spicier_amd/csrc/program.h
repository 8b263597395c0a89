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
// Value slots 0 and 1 hold the constant 1.0 (voltage-source couplings) and
// the current Gmin.  Per-iteration values live in LDS; assembly is one
// gather per pattern entry — no atomics.
#pragma once

#include <stdint.h>

namespace spx {

constexpr int kElemInts = 12;
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
  int max_row_nnz;
  const int* elems;
  const int* order;
  const int* pos_rc;   // row | col << 16, sorted by (row, col)
  const int* pos_off;  // num_pos + 1
  const int* row_ptr;  // n + 1: row r's pattern entries are [row_ptr[r], row_ptr[r+1])
  const int* contrib;  // vid << 1 | negate
  const int* rhs_off;  // n + 1
  const int* rhs_contrib;
};

// Newton / stepping configuration as the kernels read it.
struct DevNrCfg {
  double v_abstol, v_reltol, i_abstol;
  int max_iterations;
  int gmin_stepping;
  double gmin_initial, gmin_final, gmin_factor;
  int source_stepping;
  double src_initial, src_step, src_min_step;
  int precision, limiter, warm;
  int ir_max_steps;
  double ir_tol;
  double kcl_reltol, kcl_abstol, kcl_vntol;
};

// Backward-Euler transient context (transient/companion.rs:41-66): mode 0
// is the operating point the transient starts from (capacitors open, PULSE
// sources at V1, waveforms.rs:150-160); mode 1 a time step at t with step h
// (capacitor C -> G = C/h with I = G * v_prev from its negative to its
// positive node, PULSE sources at eval_pulse(t)).  vprev[e] holds the
// previous step's voltage of capacitor element e.
struct TranCtx {
  int mode;
  double t, h;
  const double* vprev;
};

// Transient run: h, step count and the probed unknowns.
struct DevTran {
  double h;
  int steps;
  int num_probes;
  const int* probes;
};

struct DevTranOut {
  double* waves;   // [B][steps + 1][num_probes]
  int* fail_step;  // [B] first step that did not converge, -1 if none
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
