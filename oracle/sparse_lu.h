// sparse_lu.h — CPU restatement of the reference's CachedSparseLu
// (spicier-solver src/linear.rs:334-409): a column order computed once from
// the first matrix's pattern (SymbolicLu::try_new), then per solve a numeric
// LU with partial pivoting on that order (Lu::try_new_with_symbolic) and the
// two triangular solves.
//
// TEST INFRASTRUCTURE ONLY (part of the oracle).
//
// faer 0.24.0's published algorithm is a COLAMD column order with supernodal
// or simplicial left-looking LU and partial pivoting.  This restatement keeps
// the same shape with simpler parts: a minimum-degree order on the pattern of
// A + A^T (lowest index on degree ties) and a simplicial left-looking
// (Gilbert-Peierls) numeric phase whose pivot is the largest |a| among rows
// not yet pivoted, the lowest row on exact ties.  Results agree with the
// dense LU to rounding, not bitwise (faer's rounding and tie-breaks are not
// reproduced).
#pragma once

#include <cstdint>
#include <vector>

namespace orc_sparse {

struct Triplet {
  int r, c;
  double v;
};

// Column-compressed matrix, duplicates summed in triplet order.
struct Csc {
  int n = 0;
  std::vector<int> colptr, row;
  std::vector<double> val;
};

Csc from_triplets(int n, const std::vector<Triplet>& t);

// The cached symbolic part: a fill-reducing column order.
struct Symbolic {
  int n = 0;
  std::vector<int> q;  // q[k] = the column eliminated at step k
};

Symbolic analyse(const Csc& a);

// Factor a (any pattern; the order comes from `sym`) and solve a x = b.
// Returns 1 (and zeros x) if a pivot column is exactly zero or x is not
// finite, else 0.
int solve(const Symbolic& sym, const Csc& a, const double* b, double* x);

}  // namespace orc_sparse
