// CPU check of the shared-symbolic sparse plan's solve descriptors and dense
// root block on N x N MNA grids (four corner-source branch rows), built with
// g++ from the library's own plan builder.  Prints one line per grid and
// exits non-zero on the first inconsistency.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sparse_plan.h"

using namespace spx;

static int check_grid(int N) {
  const int nn = N * N, n = nn + 4;
  std::vector<std::pair<int, int>> e;
  auto add = [&](int a, int b) { e.push_back({a, b}); };
  for (int i = 0; i < nn; ++i) add(i, i);
  for (int r = 0; r < N; ++r)
    for (int c = 0; c < N; ++c) {
      const int a = r * N + c;
      if (c + 1 < N) { add(a, a + 1); add(a + 1, a); }
      if (r + 1 < N) { add(a, a + N); add(a + N, a); }
    }
  const int corners[4] = {0, N - 1, nn - N, nn - 1};
  for (int k = 0; k < 4; ++k) { add(corners[k], nn + k); add(nn + k, corners[k]); }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  std::vector<int> pos;
  for (auto& p : e) pos.push_back(p.first | (p.second << 16));
  HostSparsePlan plan;
  const char* why = build_sparse_plan(n, nn, pos, plan);
  if (why && *why) { printf("grid %d: plan failed: %s\n", N, why); return 1; }
  DevSparse s = device_sparse(plan, plan.blob.data());
  const int D = s.tail;
  // descriptors: 16-byte aligned, one per position of the levels the DC
  // kernels run, each the row's own table entries
  if (plan.o_fdesc % 4 || plan.o_bkdesc % 4) { printf("grid %d: unaligned\n", N); return 1; }
  if (s.n_fdesc != s.lev_ptr[s.levels - D]) { printf("grid %d: n_fdesc\n", N); return 1; }
  for (int r = 0; r < s.n_fdesc; ++r) {
    const int i = s.lev_rows[r], nl = s.ut_ptr[i + 1] - s.ut_ptr[i];
    const int* d = s.fdesc + 4 * r;
    if (d[0] != i || d[1] != s.ut_ptr[i] || d[2] != nl || d[3] != s.diag[i] - nl) {
      printf("grid %d: fdesc %d\n", N, r);
      return 1;
    }
  }
  const int* bp = D > 0 ? s.tblev_ptr : s.blev_ptr;
  const int* br = D > 0 ? s.tblev_rows : s.blev_rows;
  const int nb = bp[D > 0 ? s.tblevels : s.blevels];
  if (s.n_bkdesc != nb) { printf("grid %d: n_bkdesc\n", N); return 1; }
  for (int r = 0; r < nb; ++r) {
    const int i = br[r];
    const int* d = s.bkdesc + 4 * r;
    if (d[0] != i || d[1] != s.col_ptr[i] || d[2] != s.col_ptr[i + 1] - s.col_ptr[i] ||
        d[3] != s.diag[i]) {
      printf("grid %d: bkdesc %d\n", N, r);
      return 1;
    }
  }
  // the staged prefix ends where the descriptors begin
  if (s.blob_ints != (int)plan.o_fdesc) { printf("grid %d: blob_ints\n", N); return 1; }
  // dense root block: rows and columns n-D .. n-1, one column per level, closed
  for (int k = 0; k < D; ++k) {
    const int L = s.levels - D + k;
    if (s.lev_ptr[L + 1] - s.lev_ptr[L] != 1 || s.lev_rows[s.lev_ptr[L]] != n - D + k) {
      printf("grid %d: tail level %d\n", N, k);
      return 1;
    }
    for (int c = s.col_ptr[n - D + k]; c < s.col_ptr[n - D + k + 1]; ++c)
      if (s.s_row[c] < n - D) { printf("grid %d: tail not closed\n", N); return 1; }
  }
  printf("grid %d: n %d nnz_lu %d levels %d tail %d descriptors %d + %d ok\n", N, n, s.nnz_lu,
         s.levels, D, s.n_fdesc, s.n_bkdesc);
  return 0;
}

int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a)
    if (check_grid(atoi(argv[a]))) return 1;
  return 0;
}
