// sparse_plan.h — the shared symbolic LU of a circuit's MNA pattern, for
// systems too large for the register LU (n > 64).
//
// Every instance of a sweep has the same sparsity pattern, so the host
// analyses it once (FaerSparseCachedBatchedSolver's symbolic-once design,
// crates/spicier-batched-sweep/src/faer_sparse_solver.rs:116-137):
//   1. ordering: minimum degree over the node unknowns of the symmetrised
//      pattern, ties to the lowest index, and each voltage-source branch
//      unknown eliminated right after the first of its nodes.  The branch
//      row's zero diagonal has then been filled (-1/G of that node), and the
//      node/branch pair acts as a Dirichlet condition on the remaining
//      nodes: a grid fed only through ideal sources keeps a nonsingular
//      Schur complement, where eliminating every node first would factor a
//      singular Laplacian and lose the solution to cancellation.  No row
//      interchanges are made; a pivot that is exactly zero marks the
//      instance singular (as nalgebra's LU does), and the FP64 KCL test
//      judges every accepted solution;
//   2. symbolic factorisation: S_k = rows below the diagonal of column k in
//      the filled pattern (equal to the columns right of it in row k);
//   3. numeric programme: for each column the LU indices of its L entries,
//      its U entries and the |S_k|^2 targets of its rank-1 update, plus the
//      transposed lists the column-oriented back substitution walks.
// The numeric factorisation on the device is right-looking Gaussian
// elimination without pivoting in this order; multipliers are formed as
// a * (1 / pivot) and updates as fma(-l, u, a).
#pragma once

#include <stdint.h>

#include <vector>

namespace spx {

struct DevSparse {
  int nnz_lu;    // values of the filled L+U pattern
  int max_s;     // max |S_k|
  int num_fill;  // LU entries that are fill (not in the assembled pattern)
  int pad;
  const int* perm;     // n: elimination step k -> unknown
  const int* diag;     // n: LU index of (k, k)
  const int* col_ptr;  // n + 1 into s_row / s_l / s_u
  const int* s_row;    // elimination index i of each entry of S_k
  const int* s_l;      // LU index of L(i, k)
  const int* s_u;      // LU index of U(k, i)
  const int* upd_ptr;  // n + 1 (64-bit safe: checked on the host) into tgt
  const int* tgt;      // LU index of (S_k[a], S_k[b]), a-major
  const int* ut_ptr;   // n + 1 into ut_row / ut_idx
  const int* ut_row;   // rows i < k with U(i, k) != 0 (elimination indices)
  const int* ut_idx;   // LU index of U(i, k)
  const int* pos_lu;   // num_pos: LU index of pattern entry q (circuit program order)
  const int* fill;     // num_fill LU indices to zero before scattering
};

// Host form of the plan: one int blob plus offsets; `build_sparse_plan`
// fills it from the circuit pattern.  Returns an empty string on success,
// else the reason the pattern cannot use the sparse path.
struct HostSparsePlan {
  int nnz_lu = 0, max_s = 0, num_fill = 0;
  int64_t num_upd = 0;
  std::vector<int> blob;
  size_t o_perm = 0, o_diag = 0, o_colptr = 0, o_srow = 0, o_sl = 0, o_su = 0, o_updptr = 0,
         o_tgt = 0, o_utptr = 0, o_utrow = 0, o_utidx = 0, o_poslu = 0, o_fill = 0;
};

// n unknowns, the first num_nodes of them node voltages; pattern entries
// (row | col << 16) sorted by (row, col).
const char* build_sparse_plan(int n, int num_nodes, const std::vector<int>& pos_rc,
                              HostSparsePlan& plan);

// Pointers of a plan whose blob sits at `base` on the device.
DevSparse device_sparse(const HostSparsePlan& plan, const int* base);

}  // namespace spx
