// sparse_plan.hip — host-side symbolic analysis for the shared-pattern LU
// (sparse_plan.h).  Runs once per circuit pattern.
#include <algorithm>
#include <functional>
#include <map>
#include <set>
#include <unordered_map>
#include <vector>

#include "sparse_plan.h"

namespace spx {

namespace {
constexpr int64_t kMaxUpdates = int64_t(1) << 24;  // 128 MB of update contributions
}

namespace {

using Adj = std::vector<std::set<int>>;

// Elimination in `order`: perm/iperm and S_k (elimination indices, sorted).
void eliminate_in(const Adj& adj0, const std::vector<int>& order, std::vector<int>& perm,
                  std::vector<int>& iperm, std::vector<std::vector<int>>& S) {
  const int n = (int)adj0.size();
  Adj adj = adj0;
  perm = order;
  iperm.assign(n, -1);
  for (int k = 0; k < n; ++k) iperm[order[k]] = k;
  S.assign(n, {});
  for (int k = 0; k < n; ++k) {
    const int v = order[k];
    std::vector<int> nb(adj[v].begin(), adj[v].end());
    for (int u : nb) S[k].push_back(iperm[u]);
    std::sort(S[k].begin(), S[k].end());
    for (int a : nb) {
      adj[a].erase(v);
      for (int b : nb)
        if (a != b) adj[a].insert(b);
    }
    adj[v].clear();
  }
}

// Branch placement after unknown v is eliminated: one unplaced branch
// unknown coupled to v both ways follows it, the one with the fewest other
// unplaced two-way neighbours (providers_of) left, and then provides in turn
// for the next, along a chain.  Each unknown provides for one branch unknown
// at most: a second one on the same node would find its diagonal cancelled
// (a node pinned by one voltage-type row leaves another attached row with a
// zero Schur complement).  `place` eliminates and marks an unknown.
template <class Place>
void place_branches(int v, const std::vector<std::vector<int>>& branches_of,
                    const std::vector<std::vector<int>>& providers_of,
                    const std::vector<char>& placed, Place&& place) {
  for (int cur = v;;) {
    int best = -1, bfree = 0;
    for (int b : branches_of[cur]) {
      if (placed[b]) continue;
      int nfree = 0;
      for (int q : providers_of[b]) nfree += !placed[q] && q != cur;
      if (best < 0 || nfree < bfree) {
        best = b;
        bfree = nfree;
      }
    }
    if (best < 0) return;
    place(best);
    cur = best;
  }
}

// Minimum degree over the plain unknowns, ties to the lowest index, the
// branch unknowns placed by place_branches (branches_of lists the branch
// unknowns each unknown is coupled to both ways, providers_of the two-way
// neighbours of each branch unknown).
std::vector<int> md_order(const Adj& adj0, const std::vector<char>& branch,
                          const std::vector<std::vector<int>>& branches_of,
                          const std::vector<std::vector<int>>& providers_of) {
  const int n = (int)adj0.size();
  Adj adj = adj0;
  std::vector<char> done(n, 0);
  std::vector<int> order;
  auto elim = [&](int v) {
    order.push_back(v);
    done[v] = 1;
    std::vector<int> nb(adj[v].begin(), adj[v].end());
    for (int a : nb) {
      adj[a].erase(v);
      for (int b : nb)
        if (a != b) adj[a].insert(b);
    }
    adj[v].clear();
  };
  int num_plain = 0;
  for (int v = 0; v < n; ++v) num_plain += !branch[v];
  for (int step = 0; step < num_plain; ++step) {
    int best = -1;
    size_t bdeg = 0;
    for (int v = 0; v < n; ++v)
      if (!done[v] && !branch[v] && (best < 0 || adj[v].size() < bdeg)) {
        best = v;
        bdeg = adj[v].size();
      }
    elim(best);
    place_branches(best, branches_of, providers_of, done, elim);
  }
  for (int v = 0; v < n; ++v)
    if (!done[v]) elim(v);
  return order;
}

// Nested dissection over the plain unknowns (recursive bisection at the
// middle breadth-first level from a pseudo-peripheral vertex), the branch
// unknowns placed by place_branches.  A path or ladder becomes a
// cyclic-reduction order: an elimination tree of depth ~2 log2 n instead of
// the chain of length n that minimum degree yields.
std::vector<int> nd_order(const Adj& adj0, const std::vector<char>& branch,
                          const std::vector<std::vector<int>>& branches_of,
                          const std::vector<std::vector<int>>& providers_of) {
  const int n = (int)adj0.size();
  std::vector<int> plain_order;
  std::vector<int> mark(n, -1), level(n, -1);
  int stamp = 0;
  auto bfs = [&](int src, int tag, std::vector<int>& seen) {
    seen.clear();
    seen.push_back(src);
    level[src] = 0;
    mark[src] = tag + 1;  // visited in this BFS
    for (size_t h = 0; h < seen.size(); ++h) {
      const int v = seen[h];
      for (int u : adj0[v])
        if (mark[u] == tag) {
          mark[u] = tag + 1;
          level[u] = level[v] + 1;
          seen.push_back(u);
        }
    }
  };
  std::function<void(std::vector<int>&)> rec = [&](std::vector<int>& V) {
    if (V.empty()) return;
    if (V.size() <= 8) {
      plain_order.insert(plain_order.end(), V.begin(), V.end());
      return;
    }
    // BFS twice from V[0] restricted to V: pseudo-peripheral start
    const int t0 = stamp;
    stamp += 4;
    for (int v : V) mark[v] = t0;
    std::vector<int> seen;
    bfs(V[0], t0, seen);
    int far = seen.back();
    std::vector<int> rest;  // other components of V
    for (int v : V)
      if (mark[v] == t0) rest.push_back(v);
    for (int v : seen) mark[v] = t0 + 2;
    bfs(far, t0 + 2, seen);
    const int nlev = level[seen.back()] + 1;
    std::vector<int> A, B, Sep;
    if (nlev >= 3) {
      std::vector<int> cnt(nlev, 0);
      for (int v : seen) cnt[level[v]]++;
      int acc = 0, sl = 0;
      for (; sl < nlev; ++sl) {
        acc += cnt[sl];
        if (2 * acc >= (int)seen.size()) break;
      }
      if (sl == 0) sl = 1;
      if (sl >= nlev - 1) sl = nlev - 2;
      for (int v : seen) (level[v] < sl ? A : level[v] > sl ? B : Sep).push_back(v);
      std::sort(A.begin(), A.end());
      std::sort(B.begin(), B.end());
      std::sort(Sep.begin(), Sep.end());
      rec(A);
      rec(B);
      plain_order.insert(plain_order.end(), Sep.begin(), Sep.end());
    } else {
      std::sort(seen.begin(), seen.end());
      plain_order.insert(plain_order.end(), seen.begin(), seen.end());
    }
    rec(rest);
  };
  std::vector<int> plain;
  for (int v = 0; v < n; ++v)
    if (!branch[v]) plain.push_back(v);
  rec(plain);
  std::vector<char> placed(n, 0);
  std::vector<int> order;
  auto put = [&](int u) {
    order.push_back(u);
    placed[u] = 1;
  };
  for (int v : plain_order) {
    put(v);
    place_branches(v, branches_of, providers_of, placed, put);
  }
  for (int v = 0; v < n; ++v)
    if (!placed[v]) order.push_back(v);
  return order;
}

// Critical path and work of an ordering, in units of one dependent update:
// every elimination level costs a fixed latency (barriers, index loads) in
// each of the factorisation and the two solves, plus its updates, which
// stream through L2 at about 16 per level-latency unit (fitted on MI355X to
// the 32x32 grid, where minimum degree wins, and the ladders, where nested
// dissection does).
double plan_cost(const std::vector<std::vector<int>>& S) {
  const int n = (int)S.size();
  std::vector<int> lev(n, 0);
  int levels = 0;
  double work = 0.0;
  for (int k = 0; k < n; ++k) {
    levels = std::max(levels, lev[k] + 1);
    for (int i : S[k]) lev[i] = std::max(lev[i], lev[k] + 1);
    work += (double)S[k].size() * (double)S[k].size() + 2.0 * S[k].size();
  }
  return 3.0 * 24.0 * levels + work / 16.0;
}

}  // namespace

const char* build_sparse_plan(int n, int num_nodes, const std::vector<int>& pos_rc,
                              HostSparsePlan& plan, const std::vector<char>* weak_diag) {
  // symmetrised adjacency without the diagonal
  Adj adj(n);
  for (int rc : pos_rc) {
    const int r = rc & 0xffff, c = rc >> 16;
    if (r != c) {
      adj[r].insert(c);
      adj[c].insert(r);
    }
  }
  // "Branch" unknowns: index >= num_nodes, no diagonal entry in the pattern
  // (an MNA voltage-source row seen through a plain matrix), or a node whose
  // diagonal holds gmin alone (weak_diag: a node between voltage sources).
  // Their diagonals are zero until an elimination fills them, so each one
  // follows an eliminated unknown it is coupled to both ways: that
  // elimination puts -a_bu a_ub / pivot_u on its diagonal.  A one-way
  // coupling (a VCVS row's control node, a CCVS row's controlling branch)
  // fills nothing and does not count.  A placed branch unknown provides for
  // the next (node, branch, node, ... along voltage sources in series).
  std::vector<char> has_diag(n, 0), branch(n, 0);
  for (int rc : pos_rc)
    if ((rc & 0xffff) == (rc >> 16)) has_diag[rc & 0xffff] = 1;
  for (int v = 0; v < n; ++v)
    branch[v] = v >= num_nodes || !has_diag[v] || (weak_diag && (*weak_diag)[v]);
  std::set<int64_t> entries;
  for (int rc : pos_rc) entries.insert(((int64_t)(rc & 0xffff) << 32) | (rc >> 16));
  auto two_way = [&](int a, int b) {
    return entries.count(((int64_t)a << 32) | b) && entries.count(((int64_t)b << 32) | a);
  };
  std::vector<std::vector<int>> branches_of(n), providers_of(n);
  for (int b = 0; b < n; ++b)
    if (branch[b])
      for (int v : adj[b])
        if (two_way(v, b)) {
          branches_of[v].push_back(b);
          providers_of[b].push_back(v);
        }
  // two orderings, the cheaper kept
  std::vector<int> perm, iperm;
  std::vector<std::vector<int>> S;
  eliminate_in(adj, md_order(adj, branch, branches_of, providers_of), perm, iperm, S);
  {
    std::vector<int> perm2, iperm2;
    std::vector<std::vector<int>> S2;
    eliminate_in(adj, nd_order(adj, branch, branches_of, providers_of), perm2, iperm2, S2);
    if (plan_cost(S2) < plan_cost(S)) {
      perm.swap(perm2);
      iperm.swap(iperm2);
      S.swap(S2);
      plan.ordering = 1;
    } else {
      plan.ordering = 0;
    }
  }
  int64_t nnz_l = 0, num_upd = 0;
  int max_s = 0;
  for (int k = 0; k < n; ++k) {
    for (int i : S[k])
      if (i <= k) return "symbolic factorisation produced an entry above the diagonal";
    nnz_l += (int64_t)S[k].size();
    num_upd += (int64_t)S[k].size() * (int64_t)S[k].size();
    max_s = std::max(max_s, (int)S[k].size());
  }
  if (num_upd > kMaxUpdates || 2 * nnz_l + n > (int64_t)0x7fffffff)
    return "the filled LU pattern is too large for the sparse sweep path";
  // row structure of L: Lrow[i] = {k < i : i in S_k}, ascending
  std::vector<std::vector<int>> Lrow(n);
  for (int k = 0; k < n; ++k)
    for (int i : S[k]) Lrow[i].push_back(k);
  // LU storage, row by row: L part, diagonal, U part (= S_i)
  std::vector<int> row_off(n + 1, 0);
  for (int i = 0; i < n; ++i) row_off[i + 1] = row_off[i] + (int)Lrow[i].size() + 1 + (int)S[i].size();
  const int nnz_lu = row_off[n];
  std::vector<std::unordered_map<int, int>> where(n);
  for (int i = 0; i < n; ++i) {
    int o = row_off[i];
    for (int j : Lrow[i]) where[i][j] = o++;
    where[i][i] = o++;
    for (int j : S[i]) where[i][j] = o++;
  }
  auto idx = [&](int i, int j) -> int {
    auto it = where[i].find(j);
    return it == where[i].end() ? -1 : it->second;
  };

  std::vector<int> diag(n), col_ptr(n + 1, 0), s_row, ut_ptr(n + 1, 0), ut_row, pos_lu, fill;
  for (int k = 0; k < n; ++k) {
    diag[k] = idx(k, k);
    for (int i : S[k]) s_row.push_back(i);
    col_ptr[k + 1] = (int)s_row.size();
    for (int i : Lrow[k]) ut_row.push_back(i);
    ut_ptr[k + 1] = (int)ut_row.size();
  }
  // forward levels: level(k) = 1 + max level(j) over j in Lrow[k]
  std::vector<int> lev(n, 0), blev(n, 0);
  int levels = 0, blevels = 0;
  for (int k = 0; k < n; ++k) {
    for (int j : Lrow[k]) lev[k] = std::max(lev[k], lev[j] + 1);
    levels = std::max(levels, lev[k] + 1);
  }
  for (int k = n - 1; k >= 0; --k) {
    for (int j : S[k]) blev[k] = std::max(blev[k], blev[j] + 1);
    blevels = std::max(blevels, blev[k] + 1);
  }
  std::vector<std::vector<int>> by_lev(levels), by_blev(blevels);
  for (int k = 0; k < n; ++k) by_lev[lev[k]].push_back(k);
  for (int k = 0; k < n; ++k) by_blev[blev[k]].push_back(k);
  std::vector<int> lev_ptr(1, 0), lev_rows, pa_ptr(1, 0), pa_l, pa_d, tb_ptr(1, 0), tb_t,
      tb_c(1, 0), cb_l, cb_u, blev_ptr(1, 0), blev_rows;
  int max_level_rows = 0;
  for (int L = 0; L < levels; ++L) {
    max_level_rows = std::max(max_level_rows, (int)by_lev[L].size());
    for (int k : by_lev[L]) {
      lev_rows.push_back(k);
      for (int i : S[k]) {
        pa_l.push_back(idx(i, k));
        pa_d.push_back(diag[k]);
      }
    }
    lev_ptr.push_back((int)lev_rows.size());
    pa_ptr.push_back((int)pa_l.size());
    // phase B: the level's targets, each with its contributions in column order
    std::map<int, std::vector<std::pair<int, int>>> contrib;
    for (int k : by_lev[L])
      for (int a : S[k])
        for (int b : S[k]) {
          const int t = idx(a, b);
          if (t < 0) return "symbolic factorisation is not closed under elimination";
          contrib[t].push_back({idx(a, k), idx(k, b)});
        }
    for (auto& kv : contrib) {
      tb_t.push_back(kv.first);
      for (auto& lu : kv.second) {
        cb_l.push_back(lu.first);
        cb_u.push_back(lu.second);
      }
      tb_c.push_back((int)cb_l.size());
    }
    tb_ptr.push_back((int)tb_t.size());
  }
  for (int L = 0; L < blevels; ++L) {
    for (int k : by_blev[L]) blev_rows.push_back(k);
    blev_ptr.push_back((int)blev_rows.size());
  }
  // narrow levels: serial cost (operations one thread would do) small enough
  // that doing them in order beats two workgroup barriers
  constexpr int kNarrow = 12;
  constexpr int kTailMin = 24, kTailMax = 48;
  std::vector<int> lev_flags(levels, 0), blev_flags(blevels, 0);
  for (int L = 0; L < levels; ++L) {
    const int fa = (pa_ptr[L + 1] - pa_ptr[L]) + (tb_c[tb_ptr[L + 1]] - tb_c[tb_ptr[L]]);
    int fw = 0;
    for (int r = lev_ptr[L]; r < lev_ptr[L + 1]; ++r) fw += 1 + (int)Lrow[lev_rows[r]].size();
    lev_flags[L] = (fa <= kNarrow ? 1 : 0) | (fw <= kNarrow ? 2 : 0);
  }
  for (int L = 0; L < blevels; ++L) {
    int bw = 0;
    for (int r = blev_ptr[L]; r < blev_ptr[L + 1]; ++r) bw += 1 + (int)S[blev_rows[r]].size();
    blev_flags[L] = bw <= kNarrow ? 1 : 0;
  }
  // dense tail: the last D levels one column each, those columns n-D .. n-1
  // in order.  The block is then closed under elimination without a check:
  // S_k lies right of k, so a tail column's entries are tail rows, and every
  // other column sits in an earlier level, its updates applied first.
  int tail = 0;
  while (tail < kTailMax && tail < levels && by_lev[levels - 1 - tail].size() == 1 &&
         by_lev[levels - 1 - tail][0] == n - 1 - tail)
    ++tail;
  if (tail < kTailMin) tail = 0;
  const int t0 = n - tail;
  std::vector<int> td_idx, tl_cnt, tblev_ptr(1, 0), tblev_rows, tblev_flags;
  int tblevels = 0;
  if (tail > 0) {
    for (int r = 0; r < tail; ++r)
      for (int c = 0; c < tail; ++c) td_idx.push_back(idx(t0 + r, t0 + c));
    for (int r = 0; r < tail; ++r) {
      int cnt = 0;
      for (int j : Lrow[t0 + r]) cnt += j < t0 ? 1 : 0;  // Lrow ascending: a prefix
      tl_cnt.push_back(cnt);
    }
    std::vector<int> tbl(n, 0);
    for (int k = t0 - 1; k >= 0; --k) {
      for (int j : S[k])
        if (j < t0) tbl[k] = std::max(tbl[k], tbl[j] + 1);
      tblevels = std::max(tblevels, tbl[k] + 1);
    }
    std::vector<std::vector<int>> by_tbl(tblevels);
    for (int k = 0; k < t0; ++k) by_tbl[tbl[k]].push_back(k);
    for (int L = 0; L < tblevels; ++L) {
      int bw = 0;
      for (int k : by_tbl[L]) {
        tblev_rows.push_back(k);
        bw += 1 + (int)S[k].size();
      }
      tblev_ptr.push_back((int)tblev_rows.size());
      tblev_flags.push_back(bw <= kNarrow ? 1 : 0);
    }
  }
  std::vector<char> covered(nnz_lu, 0);
  for (int rc : pos_rc) {
    const int t = idx(iperm[rc & 0xffff], iperm[rc >> 16]);
    if (t < 0) return "pattern entry missing from the filled pattern";
    pos_lu.push_back(t);
    covered[t] = 1;
  }
  for (int t = 0; t < nnz_lu; ++t)
    if (!covered[t]) fill.push_back(t);

  plan.nnz_lu = nnz_lu;
  plan.max_s = max_s;
  plan.num_fill = (int)fill.size();
  plan.num_upd = num_upd;
  plan.levels = levels;
  plan.blevels = blevels;
  plan.max_level_rows = max_level_rows;
  plan.tail = tail;
  plan.tblevels = tblevels;
  plan.blob.clear();
  auto put = [&](const std::vector<int>& v) {
    size_t off = plan.blob.size();
    plan.blob.insert(plan.blob.end(), v.begin(), v.end());
    while (plan.blob.size() % 4) plan.blob.push_back(0);
    return off;
  };
  plan.o_perm = put(perm);
  plan.o_diag = put(diag);
  plan.o_colptr = put(col_ptr);
  plan.o_srow = put(s_row);
  plan.o_utptr = put(ut_ptr);
  plan.o_utrow = put(ut_row);
  plan.o_levptr = put(lev_ptr);
  plan.o_levrows = put(lev_rows);
  plan.o_paptr = put(pa_ptr);
  plan.o_pal = put(pa_l);
  plan.o_pad = put(pa_d);
  plan.o_tbptr = put(tb_ptr);
  plan.o_tbt = put(tb_t);
  plan.o_tbc = put(tb_c);
  plan.o_cbl = put(cb_l);
  plan.o_cbu = put(cb_u);
  plan.o_blevptr = put(blev_ptr);
  plan.o_blevrows = put(blev_rows);
  plan.o_poslu = put(pos_lu);
  plan.o_fill = put(fill);
  plan.o_levflags = put(lev_flags);
  plan.o_blevflags = put(blev_flags);
  plan.o_tdidx = put(td_idx);
  plan.o_tlcnt = put(tl_cnt);
  plan.o_tblevptr = put(tblev_ptr);
  plan.o_tblevrows = put(tblev_rows);
  plan.o_tblevflags = put(tblev_flags);
  // the DC kernels' solve descriptors (past the prefix the transient kernels
  // stage in LDS)
  {
    std::vector<int> fd, bd;
    plan.n_fdesc = lev_ptr[levels - tail];
    for (int r = 0; r < (int)lev_rows.size(); ++r) {
      const int i = lev_rows[r], nl = ut_ptr[i + 1] - ut_ptr[i];
      fd.insert(fd.end(), {i, ut_ptr[i], nl, diag[i] - nl});
    }
    const std::vector<int>& br = tail > 0 ? tblev_rows : blev_rows;
    plan.n_bkdesc = (int)br.size();
    for (int i : br) bd.insert(bd.end(), {i, col_ptr[i], col_ptr[i + 1] - col_ptr[i], diag[i]});
    plan.o_fdesc = put(fd);
    plan.o_bkdesc = put(bd);
  }
  plan.blob.push_back(0);  // keeps every offset a valid address
  return "";
}

DevSparse device_sparse(const HostSparsePlan& p, const int* base) {
  DevSparse d;
  d.nnz_lu = p.nnz_lu;
  d.levels = p.levels;
  d.blevels = p.blevels;
  d.num_fill = p.num_fill;
  d.perm = base + p.o_perm;
  d.diag = base + p.o_diag;
  d.col_ptr = base + p.o_colptr;
  d.s_row = base + p.o_srow;
  d.ut_ptr = base + p.o_utptr;
  d.ut_row = base + p.o_utrow;
  d.lev_ptr = base + p.o_levptr;
  d.lev_rows = base + p.o_levrows;
  d.pa_ptr = base + p.o_paptr;
  d.pa_l = base + p.o_pal;
  d.pa_d = base + p.o_pad;
  d.tb_ptr = base + p.o_tbptr;
  d.tb_t = base + p.o_tbt;
  d.tb_c = base + p.o_tbc;
  d.cb_l = base + p.o_cbl;
  d.cb_u = base + p.o_cbu;
  d.blev_ptr = base + p.o_blevptr;
  d.blev_rows = base + p.o_blevrows;
  d.pos_lu = base + p.o_poslu;
  d.fill = base + p.o_fill;
  d.lev_flags = base + p.o_levflags;
  d.blev_flags = base + p.o_blevflags;
  d.tail = p.tail;
  d.tblevels = p.tblevels;
  d.td_idx = base + p.o_tdidx;
  d.tl_cnt = base + p.o_tlcnt;
  d.tblev_ptr = base + p.o_tblevptr;
  d.tblev_rows = base + p.o_tblevrows;
  d.tblev_flags = base + p.o_tblevflags;
  d.fdesc = base + p.o_fdesc;
  d.bkdesc = base + p.o_bkdesc;
  d.n_fdesc = p.n_fdesc;
  d.n_bkdesc = p.n_bkdesc;
  d.blob = base;
  d.blob_ints = (int)p.o_fdesc;
  return d;
}

}  // namespace spx
