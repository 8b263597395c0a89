// sparse_lu.cpp — see sparse_lu.h.  TEST INFRASTRUCTURE ONLY.
#include "sparse_lu.h"

#include <algorithm>
#include <cmath>

namespace orc_sparse {

Csc from_triplets(int n, const std::vector<Triplet>& t) {
  // stable sort by (column, row): duplicates keep their push order and are
  // summed in it, as SparseColMat::try_new_from_triplets sums them
  std::vector<int> idx(t.size());
  for (size_t i = 0; i < t.size(); ++i) idx[i] = (int)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    return t[a].c != t[b].c ? t[a].c < t[b].c : t[a].r < t[b].r;
  });
  Csc m;
  m.n = n;
  m.colptr.assign(n + 1, 0);
  int last_r = -1, last_c = -1;
  for (int i : idx) {
    if (t[i].c == last_c && t[i].r == last_r) {
      m.val.back() += t[i].v;
      continue;
    }
    m.row.push_back(t[i].r);
    m.val.push_back(t[i].v);
    m.colptr[t[i].c + 1]++;
    last_r = t[i].r;
    last_c = t[i].c;
  }
  for (int c = 0; c < n; ++c) m.colptr[c + 1] += m.colptr[c];
  return m;
}

Symbolic analyse(const Csc& a) {
  // minimum degree on the elimination graph of A + A^T, rows as bitsets:
  // eliminating v joins its live neighbours into a clique (each neighbour's
  // row ORs in v's row) and removes v from them
  const int n = a.n;
  const int W = (n + 63) / 64;
  std::vector<uint64_t> g((size_t)n * W, 0);
  auto row = [&](int i) { return g.data() + (size_t)i * W; };
  for (int c = 0; c < n; ++c)
    for (int p = a.colptr[c]; p < a.colptr[c + 1]; ++p) {
      const int r = a.row[p];
      if (r != c) {
        row(r)[c >> 6] |= 1ull << (c & 63);
        row(c)[r >> 6] |= 1ull << (r & 63);
      }
    }
  std::vector<int> deg(n);
  auto count = [&](int i) {
    int d = 0;
    for (int w = 0; w < W; ++w) d += __builtin_popcountll(row(i)[w]);
    return d;
  };
  for (int i = 0; i < n; ++i) deg[i] = count(i);
  std::vector<char> done(n, 0);
  std::vector<int> nb;
  Symbolic s;
  s.n = n;
  s.q.reserve(n);
  for (int k = 0; k < n; ++k) {
    int v = -1;
    for (int i = 0; i < n; ++i)
      if (!done[i] && (v < 0 || deg[i] < deg[v])) v = i;
    done[v] = 1;
    s.q.push_back(v);
    nb.clear();
    const uint64_t* rv = row(v);
    for (int w = 0; w < W; ++w)
      for (uint64_t m = rv[w]; m; m &= m - 1) nb.push_back(w * 64 + __builtin_ctzll(m));
    for (int u : nb) {
      uint64_t* ru = row(u);
      for (int w = 0; w < W; ++w) ru[w] |= rv[w];
      ru[u >> 6] &= ~(1ull << (u & 63));
      ru[v >> 6] &= ~(1ull << (v & 63));
      deg[u] = count(u);
    }
  }
  return s;
}

int solve(const Symbolic& sym, const Csc& a, const double* b, double* x) {
  const int n = a.n;
  // L by step: rows (original indices) and values below the pivot; U by step:
  // earlier steps and values above the diagonal, udiag the pivot
  std::vector<int> lptr(1, 0), lrow, uptr(1, 0), urow;
  std::vector<double> lval, uval, udiag(n);
  std::vector<int> pinv(n, -1), prow(n);
  std::vector<double> w(n, 0.0);
  std::vector<int> mark(n, -1), stack, topo, cand;
  std::vector<size_t> next;
  bool singular = false;
  for (int k = 0; k < n && !singular; ++k) {
    const int j = sym.q[k];
    // reach of A(:, j) through the columns of L (depth-first, post order)
    topo.clear();
    cand.clear();
    for (int p = a.colptr[j]; p < a.colptr[j + 1]; ++p) {
      const int r0 = a.row[p];
      w[r0] += a.val[p];
      if (mark[r0] == k) continue;
      mark[r0] = k;
      if (pinv[r0] < 0) {
        cand.push_back(r0);
        continue;
      }
      stack.assign(1, r0);
      next.assign(1, 0);
      while (!stack.empty()) {
        const int r = stack.back();
        const int s = pinv[r];
        size_t& it = next.back();
        const int end = lptr[s + 1] - lptr[s];
        bool pushed = false;
        while ((int)it < end) {
          const int rr = lrow[lptr[s] + it];
          ++it;
          if (mark[rr] == k) continue;
          mark[rr] = k;
          if (pinv[rr] < 0) {
            cand.push_back(rr);
            continue;
          }
          stack.push_back(rr);
          next.push_back(0);
          pushed = true;
          break;
        }
        if (!pushed) {
          topo.push_back(r);
          stack.pop_back();
          next.pop_back();
        }
      }
    }
    // sparse triangular solve in topological order (reverse post order)
    for (int t = (int)topo.size() - 1; t >= 0; --t) {
      const int r = topo[t];
      const int s = pinv[r];
      const double xs = w[r];
      for (int p = lptr[s]; p < lptr[s + 1]; ++p) w[lrow[p]] = (-lval[p]) * xs + w[lrow[p]];
    }
    // pivot: the largest |w| among rows not yet pivoted, lowest row on ties
    int piv = -1;
    double best = 0.0;
    for (int r : cand) {
      const double v = std::fabs(w[r]);
      if (v > best || (v == best && piv >= 0 && v > 0.0 && r < piv)) {
        best = v;
        piv = r;
      }
    }
    if (!(best > 0.0)) {
      singular = true;
      break;
    }
    const double d = w[piv];
    udiag[k] = d;
    pinv[piv] = k;
    prow[k] = piv;
    for (int r : topo) {
      urow.push_back(pinv[r]);
      uval.push_back(w[r]);
      w[r] = 0.0;
    }
    uptr.push_back((int)urow.size());
    const double inv = 1.0 / d;
    for (int r : cand) {
      if (r != piv) {
        lrow.push_back(r);
        lval.push_back(w[r] * inv);
      }
      w[r] = 0.0;
    }
    lptr.push_back((int)lrow.size());
  }
  if (singular) {
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    return 1;
  }
  // L y = P b, then U z = y, x[q[k]] = z[k]
  std::vector<double> y(b, b + n), z(n);
  for (int k = 0; k < n; ++k) {
    const double yk = y[prow[k]];
    z[k] = yk;
    for (int p = lptr[k]; p < lptr[k + 1]; ++p) y[lrow[p]] = (-lval[p]) * yk + y[lrow[p]];
  }
  for (int k = n - 1; k >= 0; --k) {
    const double zk = z[k] / udiag[k];
    z[k] = zk;
    for (int p = uptr[k]; p < uptr[k + 1]; ++p) z[urow[p]] = (-uval[p]) * zk + z[urow[p]];
  }
  for (int k = 0; k < n; ++k) x[sym.q[k]] = z[k];
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(x[i])) {
      for (int t = 0; t < n; ++t) x[t] = 0.0;
      return 1;
    }
  return 0;
}

}  // namespace orc_sparse
