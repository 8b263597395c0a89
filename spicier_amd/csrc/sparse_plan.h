// sparse_plan.h — the shared symbolic LU of a circuit's MNA pattern, for
// systems too large for the register LU (n > 64).
//
// Every instance of a sweep has the same sparsity pattern, so the host
// analyses it once (FaerSparseCachedBatchedSolver's symbolic-once design,
// crates/spicier-batched-sweep/src/faer_sparse_solver.rs:116-137):
//   1. ordering: minimum degree or nested dissection (recursive
//      breadth-first bisection) over the node unknowns of the symmetrised
//      pattern, whichever a cost model of critical path (elimination levels)
//      and update work rates cheaper — a grid keeps minimum degree, a
//      ladder takes nested dissection (16 levels instead of a 256-long
//      chain) — and each voltage-source branch unknown (an unknown past
//      num_nodes, or one without a diagonal entry in the pattern)
//      eliminated right after the first of its nodes.  The branch
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
//   3. numeric programme, level-scheduled on the elimination tree
//      (level(k) = 1 + max level of the columns that update column k):
//      per level, phase A forms the multipliers of every column in the
//      level, then phase B applies the level's rank-1 updates target by
//      target, each target summing its contributions in column order — no
//      atomics, and the result is deterministic.  The triangular solves are
//      row-oriented over the same levels (forward) and over levels counted
//      from the root (backward).  A 32x32 grid has 1,028 columns in 118
//      levels.  A level whose work is a handful of operations (a chain in
//      the tree: a ladder's is all chain) is marked narrow; a run of narrow
//      levels is done by one thread in order, with no barrier inside.
//   4. dense tail: when the last D >= 24 levels (D <= 48) are one column
//      each and those columns are the last D of the order, the trailing
//      D x D block is dense (the root separator of a grid: 48 columns for a
//      32x32 grid) and closed — every entry below or right of a tail pivot
//      lies in the block.  Kernels that take it factor the block in one
//      wave's registers after the sparse levels (row r of the block in lane
//      r, pivot row broadcast by v_readlane: the same multipliers and the
//      same fma(-l, u, a) updates in the same column order as the levels
//      would apply) and run the block's triangular solves there too, in
//      place of D levels of the factorisation and of each solve; the tail
//      rows' terms in earlier columns (a prefix of each L row) are summed
//      first.  The backward levels of the remaining rows are then counted
//      with the tail already solved (tblev_*).
// The numeric factorisation is Gaussian elimination without pivoting in
// this order; multipliers are formed as a * (1 / pivot) and updates as
// fma(-l, u, a).
#pragma once

#include <stdint.h>

#include <vector>

namespace spx {

struct DevSparse {
  int nnz_lu;    // values of the filled L+U pattern
  int levels;    // elimination-tree levels (forward)
  int blevels;   // levels counted from the root (backward solve)
  int num_fill;  // LU entries that are fill (not in the assembled pattern)
  const int* perm;     // n: elimination step k -> unknown
  const int* diag;     // n: LU index of (k, k); row k's L part ends there, its U part follows
  const int* col_ptr;  // n + 1 into s_row: S_k
  const int* s_row;    // S_k, ascending (= the U columns of row k)
  const int* ut_ptr;   // n + 1 into ut_row: L row structure of row k
  const int* ut_row;   // columns j < k with L(k, j) != 0, ascending
  const int* lev_ptr;  // levels + 1 into lev_rows
  const int* lev_rows; // rows/columns by forward level, ascending within a level
  const int* pa_ptr;   // levels + 1 into pa_l / pa_d: phase A entries of a level
  const int* pa_l;     // LU index of L(i, k)
  const int* pa_d;     // LU index of (k, k)
  const int* tb_ptr;   // levels + 1 into tb_t: phase B targets of a level
  const int* tb_t;     // LU index of the target
  const int* tb_c;     // (number of targets) + 1 into cb_l / cb_u
  const int* cb_l;     // contribution multiplier L(a, k), column order
  const int* cb_u;     // contribution U(k, b)
  const int* blev_ptr; // blevels + 1 into blev_rows
  const int* blev_rows;
  const int* lev_flags;   // levels: bit 0 the level's factor work is narrow, bit 1 its forward rows
  const int* blev_flags;  // blevels: bit 0 the level's backward rows are narrow
  const int* pos_lu;   // num_pos: LU index of pattern entry q (circuit program order)
  const int* fill;     // num_fill LU indices to zero before scattering
  int tail;            // D <= 48: dense trailing block size (0: none); rows n-D .. n-1
  int tblevels;        // backward levels of rows 0 .. n-D-1 with the tail solved
  const int* td_idx;   // D*D: LU index of block entry (n-D+r, n-D+c), -1 if not stored
  const int* tl_cnt;   // D: L terms of tail row r in columns < n-D (a prefix of its L row)
  const int* tblev_ptr;    // tblevels + 1 into tblev_rows
  const int* tblev_rows;
  const int* tblev_flags;  // bit 0: the level's rows are narrow
  // Row descriptors of the DC kernels' solves, four ints per position, after
  // the staged prefix of the blob: fdesc[r] = {i, ut_ptr[i], L terms of row
  // i, LU index of its first L term} for i = lev_rows[r]; bkdesc[r] = {i,
  // col_ptr[i], U terms right of the diagonal, diag[i]} for the r-th row of
  // the backward levels the DC kernels run (tblev_rows with a dense tail,
  // else blev_rows).  A level's rows are then one load each.
  const int* fdesc;
  const int* bkdesc;
  int n_fdesc;   // positions of the forward levels before the dense tail
  int n_bkdesc;  // positions of the DC kernels' backward levels
  const int* blob;     // the plan's one int block: every pointer above points into it
  int blob_ints;       // the prefix the transient kernels stage in LDS (excludes the descriptors)
};

// The plan with every table moved from `blob` to `to` (a copy of the block,
// e.g. in LDS).
__host__ __device__ inline DevSparse rebase_sparse(const DevSparse& s, const int* to) {
  DevSparse d = s;
  auto mv = [&](const int* p) { return to + (p - s.blob); };
  d.perm = mv(s.perm);
  d.diag = mv(s.diag);
  d.col_ptr = mv(s.col_ptr);
  d.s_row = mv(s.s_row);
  d.ut_ptr = mv(s.ut_ptr);
  d.ut_row = mv(s.ut_row);
  d.lev_ptr = mv(s.lev_ptr);
  d.lev_rows = mv(s.lev_rows);
  d.pa_ptr = mv(s.pa_ptr);
  d.pa_l = mv(s.pa_l);
  d.pa_d = mv(s.pa_d);
  d.tb_ptr = mv(s.tb_ptr);
  d.tb_t = mv(s.tb_t);
  d.tb_c = mv(s.tb_c);
  d.cb_l = mv(s.cb_l);
  d.cb_u = mv(s.cb_u);
  d.blev_ptr = mv(s.blev_ptr);
  d.blev_rows = mv(s.blev_rows);
  d.lev_flags = mv(s.lev_flags);
  d.blev_flags = mv(s.blev_flags);
  d.pos_lu = mv(s.pos_lu);
  d.fill = mv(s.fill);
  d.td_idx = mv(s.td_idx);
  d.tl_cnt = mv(s.tl_cnt);
  d.tblev_ptr = mv(s.tblev_ptr);
  d.tblev_rows = mv(s.tblev_rows);
  d.tblev_flags = mv(s.tblev_flags);
  // the descriptors lie past the staged prefix: they stay where they are
  d.blob = to;
  return d;
}

// Host form of the plan: one int blob plus offsets; `build_sparse_plan`
// fills it from the circuit pattern.  Returns an empty string on success,
// else the reason the pattern cannot use the sparse path.
struct HostSparsePlan {
  int nnz_lu = 0, max_s = 0, num_fill = 0, levels = 0, blevels = 0, max_level_rows = 0;
  int tail = 0, tblevels = 0;
  int ordering = 0;  // 0 minimum degree, 1 nested dissection
  int64_t num_upd = 0;
  std::vector<int> blob;
  size_t o_perm = 0, o_diag = 0, o_colptr = 0, o_srow = 0, o_utptr = 0, o_utrow = 0,
         o_levptr = 0, o_levrows = 0, o_paptr = 0, o_pal = 0, o_pad = 0, o_tbptr = 0,
         o_tbt = 0, o_tbc = 0, o_cbl = 0, o_cbu = 0, o_blevptr = 0, o_blevrows = 0,
         o_poslu = 0, o_fill = 0, o_levflags = 0, o_blevflags = 0, o_tdidx = 0, o_tlcnt = 0,
         o_tblevptr = 0, o_tblevrows = 0, o_tblevflags = 0, o_fdesc = 0, o_bkdesc = 0;
  int n_fdesc = 0, n_bkdesc = 0;
};

// n unknowns, the first num_nodes of them node voltages; pattern entries
// (row | col << 16) sorted by (row, col).
// weak_diag (optional, num_nodes flags): nodes whose diagonal carries gmin
// alone, ordered like branch unknowns.
const char* build_sparse_plan(int n, int num_nodes, const std::vector<int>& pos_rc,
                              HostSparsePlan& plan,
                              const std::vector<char>* weak_diag = nullptr);

// Pointers of a plan whose blob sits at `base` on the device.
DevSparse device_sparse(const HostSparsePlan& plan, const int* base);

}  // namespace spx
