// Host construction of the tree-decomposition label index.
//
// Method (the reference's TreeCount variant, src/TreeCount/TreeDecomp.cpp):
//  1. Vertex elimination (Reduce, TreeDecomp.cpp:54-208).  Eliminating x joins
//     every pair of its live neighbours (u, v) by a shortcut of length
//     l(u,x) + l(x,v); the shortcut keeps the minimum length, the number of
//     shortest restricted paths (counts multiply along x, add on ties) and the
//     maximum hop count among them (hops add along x, max on ties).  The
//     min-degree policy picks the smallest (degree, fill-degree, id) with the
//     reference's lazy re-keying of changed vertices (TreeDecomp.cpp:65-112).
//     An explicit order (order file, buildIndex.cpp:4-111) may replace it.
//  2. Tree (MakeTree, TreeDecomp.cpp:210-238): the last eliminated vertex is
//     the root, every other vertex hangs below the bag neighbour eliminated
//     first after it (GetParentNode, TreeDecomp.cpp:630-664).
//  3. Labels (makeIndexDFS, TreeDecomp.cpp:355-477): for node p and ancestor
//     a_j at depth j, dis_p[j] = min over bag vertices x at depth k >= j of
//     l(p,x) (k == j) or l(p,x) + dis_x[j] (k > j), with counts multiplied
//     and summed over ties and hops added and maxed over ties.
// Layout differs from the reference on purpose: nodes are renumbered in DFS
// preorder, rows are one flat SoA (dis / cnt / hop) with an extra self entry,
// counts are uint64 with overflow detection instead of uint32, and labels are
// computed level by level (all nodes of one depth in parallel) instead of by a
// recursive DFS, since a row depends only on its ancestors' rows.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <queue>
#include <set>
#include <tuple>

#include "bcx.h"
#include "bcx_host.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace bcx {

namespace {

struct Arc {
  int32_t nb;
  int32_t len;
  uint64_t cnt;
  uint32_t hop;
};

// Saturating count arithmetic: UINT64_MAX marks "overflowed"; it stays
// UINT64_MAX under products with nonzero counts and under sums.
inline uint64_t mul_sat(uint64_t a, uint64_t b, int32_t &ovf) {
  uint64_t r;
  if (__builtin_mul_overflow(a, b, &r)) {
    ovf = 1;
    return UINT64_MAX;
  }
  return r;
}
inline uint64_t add_sat(uint64_t a, uint64_t b, int32_t &ovf) {
  uint64_t r;
  if (__builtin_add_overflow(a, b, &r)) {
    ovf = 1;
    return UINT64_MAX;
  }
  return r;
}

// Elimination; returns elimination sequence and bags (sorted by vertex id).
void eliminate(const Graph &g, const std::vector<int32_t> &order,
               std::vector<int32_t> &elim, std::vector<std::vector<Arc>> &bags,
               int32_t &width, int32_t &ovf) {
  const int32_t n = g.n;
  std::vector<std::vector<Arc>> adj((size_t)n);
  for (int32_t x = 0; x < n; ++x) {
    auto &a = adj[(size_t)x];
    for (int64_t k = g.row[(size_t)x]; k < g.row[(size_t)x + 1]; ++k)
      a.push_back(Arc{g.col[(size_t)k], std::min(g.w[(size_t)k], kInf), 1, 1});
    std::sort(a.begin(), a.end(), [](const Arc &p, const Arc &q) { return p.nb < q.nb; });
  }
  elim.clear();
  elim.reserve((size_t)n);
  bags.assign((size_t)n, {});
  width = 0;
  std::vector<char> alive((size_t)n, 1);
  std::vector<Arc> merged;  // scratch for one neighbour's new adjacency list

  const bool min_degree = order.empty();
  std::vector<int32_t> DD, DD2, tDD, tDD2;
  std::vector<char> changed;
  // Least (DD, DD2, id) first, as the reference's ordered set: a lazy binary
  // heap whose entries that no longer match their vertex's keys (re-keyed or
  // eliminated) are dropped when they surface, so the valid top is the least
  // live (DD, DD2, id).
  using Key = std::tuple<int32_t, int32_t, int32_t>;
  std::priority_queue<Key, std::vector<Key>, std::greater<Key>> pq;
  std::vector<int32_t> seq;
  if (min_degree) {
    DD.resize((size_t)n);
    DD2.resize((size_t)n);
    tDD.resize((size_t)n);
    tDD2.resize((size_t)n);
    changed.assign((size_t)n, 0);
    for (int32_t x = 0; x < n; ++x) {
      DD[(size_t)x] = DD2[(size_t)x] = tDD[(size_t)x] = tDD2[(size_t)x] =
          (int32_t)adj[(size_t)x].size();
      pq.push(Key{DD[(size_t)x], DD2[(size_t)x], x});
    }
  } else {
    seq.assign((size_t)n, -1);
    for (int32_t v = 0; v < n; ++v) seq[(size_t)order[(size_t)v]] = v;
  }

  for (int32_t step = 0; step < n; ++step) {
    int32_t x;
    if (min_degree) {
      for (;;) {
        const Key top = pq.top();
        x = std::get<2>(top);
        if (!alive[(size_t)x] || std::get<0>(top) != DD[(size_t)x] || std::get<1>(top) != DD2[(size_t)x]) {
          pq.pop();  // stale
          continue;
        }
        if (changed[(size_t)x]) {  // lazy re-key, TreeDecomp.cpp:103-112
          pq.pop();
          DD[(size_t)x] = tDD[(size_t)x];
          DD2[(size_t)x] = tDD2[(size_t)x];
          changed[(size_t)x] = 0;
          pq.push(Key{DD[(size_t)x], DD2[(size_t)x], x});
          continue;
        }
        pq.pop();
        break;
      }
    } else {
      x = seq[(size_t)step];
    }
    elim.push_back(x);
    alive[(size_t)x] = 0;
    std::vector<Arc> nb;
    nb.swap(adj[(size_t)x]);  // only live neighbours remain in adjacency lists
    // Each neighbour y loses its arc to x and gains, towards every other
    // neighbour z, the shortcut through x: length l(y,x) + l(x,z) (kept at
    // most kInf: a shortcut that long is on no shortest path, every distance
    // staying below kInf as checked after the labels, and the clamp keeps
    // every sum of two lengths in int32), count c(y,x) c(x,z), hop sum.  An
    // existing arc y-z keeps the shorter length, adding counts and keeping
    // the larger hop on a tie (Reduce, TreeDecomp.cpp:157-193).  Both lists
    // are sorted by vertex id, so y's new list is one merge of the two: the
    // same arcs as one pair at a time with searches and inserts, in time
    // linear in the lists (the wide top of the tree has neighbour lists of
    // hundreds).  Each unordered pair is updated from both of its ends with
    // the same values, so the adjacency stays symmetric.
    // Neighbour i's new list.  The neighbours' lists are independent of each
    // other, so a large neighbourhood (the dense top of the tree) merges
    // them in parallel; each list is the same either way.
    auto merge_one = [&](size_t i, std::vector<Arc> &out, int32_t &of) {
      const Arc &a = nb[i];
      auto &ay = adj[(size_t)a.nb];
      out.clear();
      out.reserve(ay.size() + nb.size());
      size_t p = 0;
      int32_t added = 0;
      for (size_t j = 0; j <= nb.size(); ++j) {
        const int32_t z = j < nb.size() ? nb[j].nb : INT32_MAX;
        for (; p < ay.size() && ay[p].nb < z; ++p)
          if (ay[p].nb != x) out.push_back(ay[p]);
        if (j == nb.size()) break;
        if (j == i) continue;
        const Arc &b = nb[j];
        const int32_t L = (int32_t)std::min<int64_t>((int64_t)a.len + b.len, kInf);
        const uint64_t c = mul_sat(a.cnt, b.cnt, of);
        const uint32_t h = a.hop + b.hop;
        if (p < ay.size() && ay[p].nb == z) {
          Arc e = ay[p++];
          if (e.len == L) {
            e.cnt = add_sat(e.cnt, c, of);
            if (h > e.hop) e.hop = h;
          } else if (e.len > L) {
            e.len = L;
            e.cnt = c;
            e.hop = h;
          }
          out.push_back(e);
        } else {
          out.push_back(Arc{z, L, c, h});
          ++added;
        }
      }
      ay.swap(out);
      if (min_degree) {
        tDD[(size_t)a.nb] = (int32_t)ay.size();
        tDD2[(size_t)a.nb] += added;
        changed[(size_t)a.nb] = 1;
      }
    };
    int64_t work = 0;
    for (const Arc &a : nb) work += (int64_t)adj[(size_t)a.nb].size() + (int64_t)nb.size();
    if (nb.size() >= 16 && work >= (int64_t)1 << 16) {
      int32_t of_any = 0;
#pragma omp parallel reduction(| : of_any)
      {
        std::vector<Arc> out;
#pragma omp for schedule(dynamic, 2)
        for (int64_t i = 0; i < (int64_t)nb.size(); ++i) merge_one((size_t)i, out, of_any);
      }
      ovf |= of_any;
    } else {
      for (size_t i = 0; i < nb.size(); ++i) merge_one(i, merged, ovf);
    }
    if ((int32_t)nb.size() > width) width = (int32_t)nb.size();
    bags[(size_t)x] = std::move(nb);
  }
}

}  // namespace

// Every shortest distance must stay below the label INF (999999999), which
// also keeps it below 2^30 for the bag sweep's flag bit.  Entry 0 of a row is
// the label towards the root, whose restriction (vertices eliminated before
// the root) is no restriction: the true distance d(p, root).  Since
// d(s, t) <= d(s, root) + d(root, t), twice the largest such entry bounds every
// distance; graphs where it reaches the INF are refused.
void check_distance_bound(const Index &ix) {
  int64_t far = 0;
  for (int32_t p = 1; p < ix.n; ++p) far = std::max<int64_t>(far, ix.dis[(size_t)ix.row_off[(size_t)p]]);
  if (2 * far >= kInf)
    throw Error{BCX_EGRAPH, "shortest distances may reach the label INF 999999999 (farthest vertex from the "
                            "tree root at " + std::to_string(far) + ")"};
}

void need_host_labels(const Index &ix, const char *what) {
  if (!ix.host_labels)
    throw Error{BCX_EINVAL, std::string(what) +
                                " needs the label rows on the host; this index leaves them to the device"};
}

void build_index(const Graph &g, const std::vector<int32_t> &order, Index &ix, bool host_labels) {
  auto t0 = std::chrono::steady_clock::now();
  const int32_t n = g.n;
  if (!order.empty() && (int32_t)order.size() != n)
    throw Error{BCX_EINVAL, "order has wrong length"};
  ix = Index();
  ix.n = n;
  std::vector<int32_t> elim;
  std::vector<std::vector<Arc>> bags;
  eliminate(g, order, elim, bags, ix.tree_width, ix.count_overflow);

  // Tree: creation position tpos (root = 0), parent = bag vertex of max tpos.
  std::vector<int32_t> tpos((size_t)n), tvert((size_t)n), tparent((size_t)n, -1);
  std::vector<std::vector<int32_t>> children((size_t)n);
  for (int32_t k = n - 1, t = 0; k >= 0; --k, ++t) {
    int32_t v = elim[(size_t)k];
    tpos[(size_t)v] = t;
    tvert[(size_t)t] = v;
    if (k == n - 1) continue;
    const auto &b = bags[(size_t)v];
    if (b.empty())
      throw Error{BCX_EGRAPH, "graph is disconnected (vertex " + std::to_string(v + 1) +
                                  " ends a component); the index needs a connected graph"};
    int32_t best = b[0].nb;
    for (const Arc &a : b)
      if (tpos[(size_t)a.nb] > tpos[(size_t)best]) best = a.nb;
    tparent[(size_t)t] = tpos[(size_t)best];
    children[(size_t)tpos[(size_t)best]].push_back(t);
  }
  // DFS preorder over creation positions, children in creation order.
  ix.perm.resize((size_t)n);
  ix.iperm.resize((size_t)n);
  ix.depth.resize((size_t)n);
  ix.parent.resize((size_t)n);
  ix.subtree.assign((size_t)n, 1);
  std::vector<int32_t> pre_of_t((size_t)n);
  {
    std::vector<std::pair<int32_t, int32_t>> st;  // (tree pos, depth)
    st.emplace_back(0, 0);
    int32_t next = 0;
    while (!st.empty()) {
      auto [t, d] = st.back();
      st.pop_back();
      int32_t p = next++;
      pre_of_t[(size_t)t] = p;
      ix.perm[(size_t)p] = tvert[(size_t)t];
      ix.depth[(size_t)p] = d;
      const auto &ch = children[(size_t)t];
      for (size_t i = ch.size(); i-- > 0;) st.emplace_back(ch[i], d + 1);
    }
  }
  for (int32_t p = 0; p < n; ++p) ix.iperm[(size_t)ix.perm[(size_t)p]] = p;
  ix.tree_height = 0;
  for (int32_t p = 0; p < n; ++p) {
    int32_t t = tpos[(size_t)ix.perm[(size_t)p]];
    ix.parent[(size_t)p] = tparent[(size_t)t] < 0 ? -1 : pre_of_t[(size_t)tparent[(size_t)t]];
    ix.tree_height = std::max(ix.tree_height, ix.depth[(size_t)p] + 1);
  }
  for (int32_t p = n - 1; p > 0; --p) ix.subtree[(size_t)ix.parent[(size_t)p]] += ix.subtree[(size_t)p];

  // Bags in preorder ids.
  ix.bag_off.assign((size_t)n + 1, 0);
  for (int32_t p = 0; p < n; ++p)
    ix.bag_off[(size_t)p + 1] = ix.bag_off[(size_t)p] + (int64_t)bags[(size_t)ix.perm[(size_t)p]].size();
  const size_t B = (size_t)ix.bag_off[(size_t)n];
  ix.bag.resize(B);
  ix.bag_len.resize(B);
  ix.bag_cnt.resize(B);
  ix.bag_hop.resize(B);
  for (int32_t p = 0; p < n; ++p) {
    // Bag entries ordered by ancestor depth (root side first).
    std::vector<Arc> b = bags[(size_t)ix.perm[(size_t)p]];
    std::sort(b.begin(), b.end(), [&](const Arc &x, const Arc &y) {
      return ix.depth[(size_t)ix.iperm[(size_t)x.nb]] < ix.depth[(size_t)ix.iperm[(size_t)y.nb]];
    });
    for (size_t i = 0; i < b.size(); ++i) {
      const size_t e = (size_t)ix.bag_off[(size_t)p] + i;
      ix.bag[e] = ix.iperm[(size_t)b[i].nb];
      ix.bag_len[e] = b[i].len;
      ix.bag_cnt[e] = b[i].cnt;
      ix.bag_hop[e] = b[i].hop;
    }
  }

  // Label rows.
  ix.row_off.assign((size_t)n + 1, 0);
  for (int32_t p = 0; p < n; ++p)
    ix.row_off[(size_t)p + 1] = ix.row_off[(size_t)p] + ix.depth[(size_t)p] + 1;
  const int64_t E = ix.row_off[(size_t)n];
  if (!host_labels) {  // the context computes the rows (and the distance bound) on the device
    ix.host_labels = false;
    ix.fp_arcs = g.arcs();
    ix.fp_hash = graph_fingerprint(g);
    ix.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return;
  }
  ix.dis.resize((size_t)E);
  ix.cnt.resize((size_t)E);
  ix.hop.resize((size_t)E);
  std::vector<int32_t> by_depth((size_t)n), lvl_off((size_t)ix.tree_height + 1, 0);
  for (int32_t p = 0; p < n; ++p) lvl_off[(size_t)ix.depth[(size_t)p] + 1]++;
  for (int32_t d = 0; d < ix.tree_height; ++d) lvl_off[(size_t)d + 1] += lvl_off[(size_t)d];
  {
    std::vector<int32_t> f(lvl_off.begin(), lvl_off.end() - 1);
    for (int32_t p = 0; p < n; ++p) by_depth[(size_t)f[(size_t)ix.depth[(size_t)p]]++] = p;
  }
  int32_t ovf_all = ix.count_overflow;
  for (int32_t d = 0; d < ix.tree_height; ++d) {
    const int32_t lo = lvl_off[(size_t)d], hi = lvl_off[(size_t)d + 1];
#pragma omp parallel for schedule(dynamic, 16) reduction(| : ovf_all) if (hi - lo > 64)
    for (int32_t k = lo; k < hi; ++k) {
      const int32_t p = by_depth[(size_t)k];
      const int32_t D = ix.depth[(size_t)p];
      int32_t *rd = ix.dis.data() + ix.row_off[(size_t)p];
      uint64_t *rc = ix.cnt.data() + ix.row_off[(size_t)p];
      uint32_t *rh = ix.hop.data() + ix.row_off[(size_t)p];
      for (int32_t j = 0; j < D; ++j) {
        rd[j] = kInf;
        rc[j] = 0;
        rh[j] = 0;
      }
      rd[D] = 0;
      rc[D] = 1;
      rh[D] = 0;
      const auto &b = bags[(size_t)ix.perm[(size_t)p]];
      // Entry j of the row (the ancestor at depth j) is the minimum over the
      // direct shortcut, where that ancestor is in the bag, and every bag
      // entry x's shortcut plus x's own entry j; its count is the saturating
      // sum, and its hop the maximum, over the candidates attaining the
      // minimum.  Sums of saturating products and maxima do not depend on the
      // order of their terms, so two passes over the bag (the minimum first,
      // in a branch-free loop the compiler vectorizes, then counts and hops
      // from the attaining candidates) give the entries of one running
      // min-and-count pass (makeIndexDFS, TreeDecomp.cpp:355-477).  Lengths
      // and label distances are at most kInf (999999999), so a candidate sum
      // stays below 2^31; candidates through an entry at kInf (no restricted
      // path) are left out, as they never attain a distance below kInf.
      for (const Arc &a : b) rd[(size_t)ix.depth[(size_t)ix.iperm[(size_t)a.nb]]] = a.len;
      for (const Arc &a : b) {
        const int32_t x = ix.iperm[(size_t)a.nb];
        const int32_t kpos = ix.depth[(size_t)x];
        const int32_t *xd = ix.dis.data() + ix.row_off[(size_t)x];
        const int32_t len = a.len;
        for (int32_t j = 0; j < kpos; ++j) rd[j] = std::min(rd[j], len + xd[j]);
      }
      int32_t ovf = 0;
      for (const Arc &a : b) {
        const int32_t x = ix.iperm[(size_t)a.nb];
        const int32_t kpos = ix.depth[(size_t)x];
        if (rd[kpos] == a.len) {  // the direct shortcut attains it
          rc[kpos] = add_sat(rc[kpos], a.cnt, ovf);
          rh[kpos] = std::max(rh[kpos], a.hop);
        }
        const int32_t *xd = ix.dis.data() + ix.row_off[(size_t)x];
        const uint64_t *xc = ix.cnt.data() + ix.row_off[(size_t)x];
        const uint32_t *xh = ix.hop.data() + ix.row_off[(size_t)x];
        const int32_t len = a.len;
        for (int32_t j = 0; j < kpos; ++j) {
          if (len + xd[j] != rd[j] || xd[j] >= kInf) continue;
          rc[j] = add_sat(rc[j], mul_sat(a.cnt, xc[j], ovf), ovf);
          rh[j] = std::max(rh[j], a.hop + xh[j]);
        }
      }
      ovf_all |= ovf;
    }
  }
  ix.count_overflow = ovf_all;
  check_distance_bound(ix);
  ix.fp_arcs = g.arcs();
  ix.fp_hash = graph_fingerprint(g);
  ix.build_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int32_t lca_depth(const Index &ix, int32_t p, int32_t q) {
  // Ancestor test on preorder ids: a is an ancestor-or-self of x iff
  // a <= x < a + subtree[a].
  while (!(p <= q && q < p + ix.subtree[(size_t)p])) p = ix.parent[(size_t)p];
  return ix.depth[(size_t)p];
}

// Pair query over the common prefix of two rows: the aggregation of
// MHLQBC.cu:356-404 / TreeDecomp.cpp:969-1017, with the ancestor case folded
// into the self entry.
void index_query(const Index &ix, int32_t s, int32_t t, int32_t &d, uint64_t &sigma,
                 uint32_t &hop, bool &ovf) {
  const int32_t p = ix.iperm[(size_t)s], q = ix.iperm[(size_t)t];
  const int32_t L = lca_depth(ix, std::min(p, q), std::max(p, q));
  const int64_t op = ix.row_off[(size_t)p], oq = ix.row_off[(size_t)q];
  int32_t best = 2 * kInf + 1;
  uint64_t c = 0;
  uint32_t h = 0;
  int32_t of = 0;
  for (int32_t i = 0; i <= L; ++i) {
    const int32_t v = ix.dis[(size_t)(op + i)] + ix.dis[(size_t)(oq + i)];
    if (v > best) continue;
    const uint64_t cc = mul_sat(ix.cnt[(size_t)(op + i)], ix.cnt[(size_t)(oq + i)], of);
    const uint32_t hh = ix.hop[(size_t)(op + i)] + ix.hop[(size_t)(oq + i)];
    if (v < best) {
      best = v;
      c = cc;
      h = hh;
    } else {
      c = add_sat(c, cc, of);
      if (hh > h) h = hh;
    }
  }
  d = best;
  sigma = c;
  hop = h;
  // Counts saturate at UINT64_MAX and saturation propagates through products
  // and sums, so only a saturated final count means the true count overflowed
  // (a saturated candidate that loses the minimum does not matter).
  ovf = c == UINT64_MAX;
}

// ---- min-degree order file (QBC input) --------------------------------------
// The contraction order of Graph::GraphConstract (buildIndex.cpp:759-925):
// repeatedly contract a vertex of least current degree in the filled graph
// (ties to the smaller id: its secondary key compares _DD2[x] with itself,
// buildIndex.cpp:732, so it never decides), joining the contracted vertex's
// remaining neighbours pairwise.  Every neighbour whose degree changed is
// re-keyed at once (buildIndex.cpp:895-904), so the key is always the current
// degree.  Like the reference built without its parallel branch
// (buildIndex.cpp:847), a vertex with more than 1000 remaining neighbours
// adds no fill.  order[v] = contraction rank of v.
void min_degree_order(const Graph &g, std::vector<int32_t> &order, int64_t *shortcut_arcs,
                      std::vector<int32_t> *width_log) {
  const int32_t n = g.n;
  std::vector<std::vector<int32_t>> adj((size_t)n);
  for (int32_t v = 0; v < n; ++v) {
    auto &a = adj[(size_t)v];
    a.assign(g.col.begin() + g.row[(size_t)v], g.col.begin() + g.row[(size_t)v + 1]);
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  auto drop = [&](int32_t u, int32_t v) {
    auto &a = adj[(size_t)u];
    auto it = std::lower_bound(a.begin(), a.end(), v);
    if (it != a.end() && *it == v) a.erase(it);
  };
  std::vector<int32_t> key((size_t)n), merged;
  // Least (degree, id) first.  A re-keyed vertex gets a new entry; an old
  // one is skipped when it surfaces (its degree no longer matches, or the
  // vertex is gone), so the valid top is always the least live (key, id).
  std::priority_queue<int64_t, std::vector<int64_t>, std::greater<int64_t>> q;
  auto entry = [](int32_t d, int32_t v) { return ((int64_t)d << 32) | (int64_t)(uint32_t)v; };
  std::vector<char> gone((size_t)n, 0);
  for (int32_t v = 0; v < n; ++v) {
    key[(size_t)v] = (int32_t)adj[(size_t)v].size();
    q.push(entry(key[(size_t)v], v));
  }
  order.assign((size_t)n, 0);
  int64_t arcs_out = 0;
  int32_t width = 0;
  if (width_log) width_log->clear();
  for (int32_t k = 0; k < n; ++k) {
    if (width_log && k % 100000 == 0) width_log->push_back(width);
    int32_t x;
    for (;;) {
      const int64_t top = q.top();
      q.pop();
      x = (int32_t)(uint32_t)top;
      if (!gone[(size_t)x] && key[(size_t)x] == (int32_t)(top >> 32)) break;
    }
    gone[(size_t)x] = 1;
    order[(size_t)x] = k;
    const std::vector<int32_t> nb = std::move(adj[(size_t)x]);
    adj[(size_t)x].clear();
    arcs_out += (int64_t)nb.size();
    width = std::max(width, (int32_t)nb.size());
    // Each neighbour y loses x and, unless the neighbourhood is over 1000,
    // gains the other neighbours: one merge of two sorted lists per y.  The
    // neighbours' lists are independent, so a large neighbourhood merges
    // them in parallel.
    auto merge_one = [&](int32_t y, std::vector<int32_t> &out) {
      auto &a = adj[(size_t)y];
      out.clear();
      size_t i = 0, j = 0;
      while (i < a.size() || j < nb.size()) {
        int32_t z;
        if (j == nb.size() || (i < a.size() && a[i] < nb[j])) {
          z = a[i++];
        } else if (i == a.size() || nb[j] < a[i]) {
          z = nb[j++];
        } else {
          z = a[i++];
          ++j;
        }
        if (z != x && z != y) out.push_back(z);
      }
      a.swap(out);
    };
    if (nb.size() > 1000) {
      for (int32_t y : nb) drop(y, x);
    } else {
      int64_t work = 0;
      for (int32_t y : nb) work += (int64_t)adj[(size_t)y].size() + (int64_t)nb.size();
      if (nb.size() >= 16 && work >= (int64_t)1 << 16) {
#pragma omp parallel
        {
          std::vector<int32_t> out;
#pragma omp for schedule(dynamic, 2)
          for (int64_t k = 0; k < (int64_t)nb.size(); ++k) merge_one(nb[(size_t)k], out);
        }
      } else {
        for (int32_t y : nb) merge_one(y, merged);
      }
    }
    for (int32_t y : nb) {
      const int32_t d = (int32_t)adj[(size_t)y].size();
      if (d != key[(size_t)y]) {
        key[(size_t)y] = d;
        q.push(entry(d, y));
      }
    }
  }
  if (shortcut_arcs) *shortcut_arcs = arcs_out;
}

// ---- H2H distance labels (QBC path) ---------------------------------------
// Global distance labels, the top-down recurrence of makeIndexDFS
// (buildIndex.cpp:245-268) over the bags: for node p and its ancestor y at
// depth j, d(p, y) = min over bag entries (x at depth k, shortcut length w)
// of w + d(x, y), where d(x, y) is x's label at depth j when x is deeper
// (k > j), y's label at depth k when y is deeper (k < j), and 0 when x = y.
// Parents' rows are complete before their descendants', so each level runs in
// parallel.  Separators follow optimizeVS (buildIndex.cpp:313-386): a node
// with fewer than two children needs only its own depth (it is the LCA of a
// pair only when one end is the node itself); with two, the bag depths of the
// child with the smaller bag; with more, the bag depths of every child but
// the one with the most depths unique to it (any two children include one of
// the others, and a path between two child subtrees crosses both children's
// bags); in every case the node's own depth last.
namespace {

// Depth of LCA and the LCA's preorder id for preorder ids p, q.
int32_t lca_node(const Index &ix, int32_t p, int32_t q) {
  if (p > q) std::swap(p, q);
  while (!(p <= q && q < p + ix.subtree[(size_t)p])) p = ix.parent[(size_t)p];
  return p;
}

// P2H list of node s below branch node v (preorder ids), or null when v has
// no table.
const int32_t *p2h_list(const Index &ix, int32_t v, int32_t s, int64_t &len) {
  for (size_t b = 0; b < ix.p2h_node.size(); ++b)
    if (ix.p2h_node[b] == v) {
      const int64_t i = ix.p2h_base[b] + (s - v);
      len = ix.p2h_off[(size_t)i + 1] - ix.p2h_off[(size_t)i];
      return ix.p2h.data() + ix.p2h_off[(size_t)i];
    }
  return nullptr;
}

// The P2H branch tables of QBC (SVSP_P2H and P2H_3, buildIndex.cpp:565-719),
// with the reference's parameters k = 10, d = 5, R = 1e6 (QBC.cu:431) and a
// fixed seed instead of the clock (buildIndex.cpp:608):
//  1. branch nodes (two or more children) scored by the share of vertex
//     pairs whose LCA they are, 2 sum_{i<j} |T_i| |T_j| / n^2 over their
//     child subtrees; the k best (ties to the smaller vertex id);
//  2. for each, and each node s of its subtree, the separator depths c such
//     that no other separator vertex u has d(s,u) <= d(s,c) and
//     d(s,c) = d(s,u) + d(u,c): a path from s through c can go through u
//     instead, so the minimum over s's list alone is the distance;
//  3. R seeded random pairs whose LCA is one of them (neither end the LCA)
//     credit it |separator| - |shorter of the two lists|; the d best keep
//     their tables.
void build_p2h(Index &ix, int k, int d, int64_t R, uint64_t seed) {
  const int32_t n = ix.n;
  ix.p2h_node.clear();
  ix.p2h_base.clear();
  ix.p2h_off.clear();
  ix.p2h.clear();
  std::vector<int32_t> nch((size_t)n, 0);
  for (int32_t p = 1; p < n; ++p) nch[(size_t)ix.parent[(size_t)p]]++;
  std::vector<std::pair<double, int32_t>> score;  // (p, vertex id)
  {
    std::vector<double> acc((size_t)n, 0.0), sq((size_t)n, 0.0);
    for (int32_t p = 1; p < n; ++p) {
      const double t = ix.subtree[(size_t)p];
      acc[(size_t)ix.parent[(size_t)p]] += t;
      sq[(size_t)ix.parent[(size_t)p]] += t * t;
    }
    for (int32_t p = 0; p < n; ++p)
      if (nch[(size_t)p] >= 2)  // sum_{i<j} a_i a_j = ((sum a)^2 - sum a^2) / 2
        score.emplace_back((acc[(size_t)p] * acc[(size_t)p] - sq[(size_t)p]) / ((double)n * n),
                           ix.perm[(size_t)p]);
  }
  std::sort(score.begin(), score.end(), [](const auto &a, const auto &b) {
    return a.first != b.first ? a.first > b.first : a.second < b.second;
  });
  if ((int)score.size() > k) score.resize((size_t)k);
  if (score.empty()) return;
  // Tables of the candidates.
  struct Table {
    int32_t v;
    std::vector<int64_t> off;  // subtree size + 1
    std::vector<int32_t> list;
  };
  std::vector<Table> tabs(score.size());
  for (size_t b = 0; b < score.size(); ++b) {
    Table &T = tabs[b];
    const int32_t v = ix.iperm[(size_t)score[b].second];
    T.v = v;
    const int32_t size = ix.subtree[(size_t)v];
    const int64_t s0 = ix.sep_off[(size_t)v], s1 = ix.sep_off[(size_t)v + 1];
    const int32_t m = (int32_t)(s1 - s0);
    const int32_t *sepv = ix.sep.data() + s0;
    // Ancestors of v by depth, then d(u, c) between separator vertices: the
    // deeper one's label at the shallower one's depth.
    const int32_t Dv = ix.depth[(size_t)v];
    std::vector<int32_t> anc((size_t)Dv + 1);
    for (int32_t a = v, j = Dv; j >= 0; a = ix.parent[(size_t)a], --j) anc[(size_t)j] = a;
    std::vector<int64_t> duc((size_t)m * m, 0);
    for (int32_t i = 0; i < m; ++i)
      for (int32_t j = 0; j < m; ++j) {
        const int32_t lo = std::max(sepv[i], sepv[j]), hi = std::min(sepv[i], sepv[j]);
        duc[(size_t)i * m + j] = ix.gdis[(size_t)(ix.row_off[(size_t)anc[(size_t)lo]] + hi)];
      }
    std::vector<std::vector<int32_t>> lists((size_t)size);
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t r = 1; r < size; ++r) {
      const int32_t s = v + r;
      const int64_t row = ix.row_off[(size_t)s];
      auto &out = lists[(size_t)r];
      for (int32_t c = 0; c < m; ++c) {
        const int64_t dsc = ix.gdis[(size_t)(row + sepv[c])];
        bool keep = true;
        for (int32_t u = 0; u < m && keep; ++u) {
          if (u == c) continue;
          const int64_t dsu = ix.gdis[(size_t)(row + sepv[u])];
          if (dsu <= dsc && dsc == dsu + duc[(size_t)u * m + c]) keep = false;
        }
        if (keep) out.push_back(sepv[c]);
      }
    }
    T.off.assign((size_t)size + 1, 0);
    for (int32_t r = 0; r < size; ++r) T.off[(size_t)r + 1] = T.off[(size_t)r] + (int64_t)lists[(size_t)r].size();
    T.list.reserve((size_t)T.off[(size_t)size]);
    for (auto &l : lists) T.list.insert(T.list.end(), l.begin(), l.end());
  }
  // Credit each candidate over R seeded random pairs (splitmix64 stream).
  std::vector<double> gain(tabs.size(), 0.0);
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  for (int64_t i = 0; i < R; ++i) {
    const int32_t p = (int32_t)(next() % (uint64_t)n), q = (int32_t)(next() % (uint64_t)n);
    if (p == q) {
      --i;
      continue;
    }
    const int32_t a = lca_node(ix, p, q);
    if (a == p || a == q) continue;
    for (size_t b = 0; b < tabs.size(); ++b)
      if (tabs[b].v == a) {
        const auto &T = tabs[b];
        const int64_t lp = T.off[(size_t)(p - a) + 1] - T.off[(size_t)(p - a)];
        const int64_t lq = T.off[(size_t)(q - a) + 1] - T.off[(size_t)(q - a)];
        gain[b] += (double)(ix.sep_off[(size_t)a + 1] - ix.sep_off[(size_t)a]) - (double)std::min(lp, lq);
      }
  }
  std::vector<size_t> ord(tabs.size());
  for (size_t b = 0; b < ord.size(); ++b) ord[b] = b;
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return gain[a] > gain[b]; });
  if ((int)ord.size() > d) ord.resize((size_t)d);
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return tabs[a].v < tabs[b].v; });
  for (size_t b : ord) {
    const Table &T = tabs[b];
    ix.p2h_node.push_back(T.v);
    ix.p2h_base.push_back((int64_t)ix.p2h_off.size());
    const int64_t shift = (int64_t)ix.p2h.size();
    for (int64_t o : T.off) ix.p2h_off.push_back(o + shift);
    ix.p2h.insert(ix.p2h.end(), T.list.begin(), T.list.end());
  }
}

}  // namespace

void build_h2h(Index &ix) {
  const int32_t n = ix.n, H = ix.tree_height;
  ix.gdis.assign((size_t)ix.entries(), kInf);
  std::vector<int32_t> lvl_off((size_t)H + 1, 0), lvl((size_t)n);
  for (int32_t p = 0; p < n; ++p) lvl_off[(size_t)ix.depth[(size_t)p] + 1]++;
  for (int32_t k = 0; k < H; ++k) lvl_off[(size_t)k + 1] += lvl_off[(size_t)k];
  {
    std::vector<int32_t> f(lvl_off.begin(), lvl_off.end() - 1);
    for (int32_t p = 0; p < n; ++p) lvl[(size_t)f[(size_t)ix.depth[(size_t)p]]++] = p;
  }
  for (int32_t k = 0; k < H; ++k) {
#pragma omp parallel
    {
      std::vector<int32_t> anc;
#pragma omp for schedule(dynamic, 64)
      for (int32_t q = lvl_off[(size_t)k]; q < lvl_off[(size_t)k + 1]; ++q) {
        const int32_t p = lvl[(size_t)q], D = ix.depth[(size_t)p];
        anc.assign((size_t)D, 0);
        for (int32_t a = ix.parent[(size_t)p], j = D - 1; j >= 0; a = ix.parent[(size_t)a], --j)
          anc[(size_t)j] = a;
        const int64_t row = ix.row_off[(size_t)p];
        ix.gdis[(size_t)(row + D)] = 0;
        for (int32_t j = 0; j < D; ++j) {
          int64_t g = kInf;
          for (int64_t e = ix.bag_off[(size_t)p]; e < ix.bag_off[(size_t)p + 1]; ++e) {
            const int32_t x = ix.bag[(size_t)e], kx = ix.depth[(size_t)x];
            int64_t z;
            if (kx == j)
              z = 0;
            else if (kx < j)
              z = ix.gdis[(size_t)(ix.row_off[(size_t)anc[(size_t)j]] + kx)];
            else
              z = ix.gdis[(size_t)(ix.row_off[(size_t)x] + j)];
            g = std::min<int64_t>(g, z + ix.bag_len[(size_t)e]);
          }
          ix.gdis[(size_t)(row + j)] = (int32_t)std::min<int64_t>(g, kInf);
        }
      }
    }
  }
  // Children in creation (preorder) order.
  std::vector<int32_t> ch_off((size_t)n + 1, 0), ch((size_t)std::max(0, n - 1));
  for (int32_t p = 1; p < n; ++p) ch_off[(size_t)ix.parent[(size_t)p] + 1]++;
  for (int32_t p = 0; p < n; ++p) ch_off[(size_t)p + 1] += ch_off[(size_t)p];
  {
    std::vector<int32_t> f(ch_off.begin(), ch_off.end() - 1);
    for (int32_t p = 1; p < n; ++p) ch[(size_t)f[(size_t)ix.parent[(size_t)p]]++] = p;
  }
  auto bag_size = [&](int32_t c) { return ix.bag_off[(size_t)c + 1] - ix.bag_off[(size_t)c]; };
  ix.sep_off.assign((size_t)n + 1, 0);
  ix.sep.clear();
  std::vector<int32_t> count((size_t)H, 0), touched;
  for (int32_t p = 0; p < n; ++p) {
    const int32_t own = ix.depth[(size_t)p];
    const int32_t c0 = ch_off[(size_t)p], nch = ch_off[(size_t)p + 1] - c0;
    if (nch == 2) {
      const int32_t c1 = ch[(size_t)c0], c2 = ch[(size_t)c0 + 1];
      const int32_t tmp = bag_size(c1) < bag_size(c2) ? c1 : c2;
      for (int64_t e = ix.bag_off[(size_t)tmp + 1] - 1; e >= ix.bag_off[(size_t)tmp]; --e) {
        const int32_t d = ix.depth[(size_t)ix.bag[(size_t)e]];
        if (d != own) ix.sep.push_back(d);
      }
    } else if (nch > 2) {
      touched.clear();
      for (int32_t i = 0; i < nch; ++i) {
        const int32_t c = ch[(size_t)(c0 + i)];
        for (int64_t e = ix.bag_off[(size_t)c]; e < ix.bag_off[(size_t)c + 1]; ++e) {
          const int32_t d = ix.depth[(size_t)ix.bag[(size_t)e]];
          if (count[(size_t)d]++ == 0) touched.push_back(d);
        }
      }
      int32_t idx = 0, best = -1;
      for (int32_t i = 0; i < nch; ++i) {
        const int32_t c = ch[(size_t)(c0 + i)];
        int32_t a = 0;
        for (int64_t e = ix.bag_off[(size_t)c]; e < ix.bag_off[(size_t)c + 1]; ++e)
          a += count[(size_t)ix.depth[(size_t)ix.bag[(size_t)e]]] == 1;
        if (a > best) {
          best = a;
          idx = i;
        }
      }
      {
        const int32_t c = ch[(size_t)(c0 + idx)];
        for (int64_t e = ix.bag_off[(size_t)c]; e < ix.bag_off[(size_t)c + 1]; ++e)
          count[(size_t)ix.depth[(size_t)ix.bag[(size_t)e]]]--;
      }
      std::sort(touched.begin(), touched.end());
      for (int32_t d : touched) {
        if (count[(size_t)d] > 0 && d != own) ix.sep.push_back(d);
        count[(size_t)d] = 0;
      }
    }
    ix.sep.push_back(own);
    ix.sep_off[(size_t)p + 1] = (int64_t)ix.sep.size();
  }
  build_p2h(ix, 10, 5, 1000000, 0x5eed);
}

int32_t h2h_query(const Index &ix, int32_t s, int32_t t) {
  if (ix.sep_off.empty()) throw Error{BCX_EINVAL, "H2H labels not built"};
  const int32_t p = ix.iperm[(size_t)s], q = ix.iperm[(size_t)t];
  if (p == q) return 0;
  const int32_t L = lca_depth(ix, std::min(p, q), std::max(p, q));
  const int32_t dp = ix.depth[(size_t)p], dq = ix.depth[(size_t)q];
  const int64_t op = ix.row_off[(size_t)p], oq = ix.row_off[(size_t)q];
  if (L == dp) return ix.gdis[(size_t)(oq + dp)];  // p is an ancestor of q
  if (L == dq) return ix.gdis[(size_t)(op + dq)];
  int32_t a = p;
  while (ix.depth[(size_t)a] > L) a = ix.parent[(size_t)a];
  int64_t best = 2 * (int64_t)kInf;
  // A P2H branch node: the shorter of the two ends' projected lists
  // (QBC.cu:290-312); otherwise the optimizeVS separator (QBC.cu:313-329).
  int64_t lp = 0, lq = 0;
  const int32_t *Lp = p2h_list(ix, a, p, lp), *Lq = Lp ? p2h_list(ix, a, q, lq) : nullptr;
  const int32_t *list = ix.sep.data() + ix.sep_off[(size_t)a];
  int64_t len = ix.sep_off[(size_t)a + 1] - ix.sep_off[(size_t)a];
  if (Lp) {
    list = lp <= lq ? Lp : Lq;
    len = std::min(lp, lq);
  }
  for (int64_t i = 0; i < len; ++i) {
    const int32_t k = list[i];
    best = std::min<int64_t>(best, (int64_t)ix.gdis[(size_t)(op + k)] + ix.gdis[(size_t)(oq + k)]);
  }
  return (int32_t)std::min<int64_t>(best, kInf);
}

void h2h_scan_stats(const Index &ix, int64_t pairs, uint64_t seed, double &mean_sep, double &mean_p2h) {
  if (ix.sep_off.empty()) throw Error{BCX_EINVAL, "H2H labels not built"};
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  double a_sep = 0, a_p2h = 0;
  for (int64_t i = 0; i < pairs; ++i) {
    const int32_t p = (int32_t)(next() % (uint64_t)ix.n), q = (int32_t)(next() % (uint64_t)ix.n);
    const int32_t a = lca_node(ix, p, q);
    if (a == p || a == q) {  // one lookup (QBC.cu:282-288)
      a_sep += 1;
      a_p2h += 1;
      continue;
    }
    const double m = (double)(ix.sep_off[(size_t)a + 1] - ix.sep_off[(size_t)a]);
    int64_t lp = 0, lq = 0;
    const int32_t *Lp = p2h_list(ix, a, p, lp);
    if (Lp) p2h_list(ix, a, q, lq);
    a_sep += m;
    a_p2h += Lp ? (double)std::min(lp, lq) : m;
  }
  mean_sep = pairs ? a_sep / (double)pairs : 0.0;
  mean_p2h = pairs ? a_p2h / (double)pairs : 0.0;
}

// ---- tight bag pairs (bag-sweep unpacking) --------------------------------
// For every node z, in the order `nodes`, and every pair of its bag entries
// i < j (ancestors a above b), the pair is tight when the shortcut z-a plus
// z-b is as short as the shortcut a-b, which is entry e of b's bag (found by
// binary search over b's depth-ordered bag; every pair of a bag is joined by
// a fill shortcut).  Out: tz[q] = z, tab[3q..3q+2] = (e, i, j), and
// tfr[q] = c(a,z) c(z,b) / c(a,b), in (nodes order, i, j) order.  Two
// parallel passes (count, then fill at prefix offsets) keep that order.
void tight_pairs(const Index &ix, const std::vector<int32_t> &nodes, std::vector<int32_t> &tz,
                 std::vector<int32_t> &tab, std::vector<double> &tfr) {
  const int64_t N = (int64_t)nodes.size();
  std::vector<int64_t> cnt((size_t)N + 1, 0);
  int bad = 0;
  auto find = [&](int64_t i, int64_t j) -> int64_t {
    const int32_t a = ix.bag[(size_t)i], b = ix.bag[(size_t)j];
    const int32_t da = ix.depth[(size_t)a];
    int64_t lo = ix.bag_off[(size_t)b], hi = ix.bag_off[(size_t)b + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (ix.depth[(size_t)ix.bag[(size_t)mid]] < da)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo == ix.bag_off[(size_t)b + 1] || ix.bag[(size_t)lo] != a) return -1;
    return lo;
  };
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (int64_t q = 0; q < N; ++q) {
    const int32_t z = nodes[(size_t)q];
    int64_t c = 0;
    for (int64_t i = ix.bag_off[(size_t)z]; i < ix.bag_off[(size_t)z + 1]; ++i)
      for (int64_t j = i + 1; j < ix.bag_off[(size_t)z + 1]; ++j) {
        const int64_t e = find(i, j);
        if (e < 0) {
          bad = 1;
          continue;
        }
        c += (int64_t)ix.bag_len[(size_t)i] + ix.bag_len[(size_t)j] == ix.bag_len[(size_t)e];
      }
    cnt[(size_t)q + 1] = c;
  }
  if (bad) throw Error{BCX_EGRAPH, "index bag is not closed under its shortcuts"};
  for (int64_t q = 0; q < N; ++q) cnt[(size_t)q + 1] += cnt[(size_t)q];
  const int64_t P = cnt[(size_t)N];
  if (P > 0x7fffffff) throw Error{BCX_EGRAPH, "more than 2^31 tight bag pairs"};
  tz.assign((size_t)P, 0);
  tab.assign((size_t)(3 * P), 0);
  tfr.assign((size_t)P, 0.0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < N; ++q) {
    const int32_t z = nodes[(size_t)q];
    int64_t k = cnt[(size_t)q];
    for (int64_t i = ix.bag_off[(size_t)z]; i < ix.bag_off[(size_t)z + 1]; ++i)
      for (int64_t j = i + 1; j < ix.bag_off[(size_t)z + 1]; ++j) {
        const int64_t e = find(i, j);
        if ((int64_t)ix.bag_len[(size_t)i] + ix.bag_len[(size_t)j] != ix.bag_len[(size_t)e]) continue;
        tz[(size_t)k] = z;
        tab[(size_t)(3 * k)] = (int32_t)e;
        tab[(size_t)(3 * k + 1)] = (int32_t)i;
        tab[(size_t)(3 * k + 2)] = (int32_t)j;
        tfr[(size_t)k] = (double)ix.bag_cnt[(size_t)i] * (double)ix.bag_cnt[(size_t)j] / (double)ix.bag_cnt[(size_t)e];
        ++k;
      }
  }
}

// ---- binary image ---------------------------------------------------------
namespace {
const char kMagic[8] = {'B', 'C', 'X', 'I', 'D', 'X', '0', '3'};
template <class T>
void wvec(FILE *f, const std::vector<T> &v) {
  int64_t k = (int64_t)v.size();
  if (std::fwrite(&k, sizeof k, 1, f) != 1 ||
      (k && std::fwrite(v.data(), sizeof(T), (size_t)k, f) != (size_t)k))
    throw Error{BCX_EIO, "index write failed"};
}
template <class T>
void rvec(FILE *f, std::vector<T> &v, int64_t max_len) {
  int64_t k;
  if (std::fread(&k, sizeof k, 1, f) != 1 || k < 0 || k > max_len)
    throw Error{BCX_EIO, "index file corrupt"};
  v.resize((size_t)k);
  if (k && std::fread(v.data(), sizeof(T), (size_t)k, f) != (size_t)k)
    throw Error{BCX_EIO, "index file truncated"};
}
}  // namespace

// Checks the H2H labels, separators and P2H tables of an index whose tree
// is valid (iperm and subtree derived); throws BCX_EIO naming `what`.
void check_h2h(const Index &ix, const std::string &what) {
  auto bad = [&](const std::string &m) { throw Error{BCX_EIO, what + ": " + m}; };
  const int32_t n = ix.n;
  const int64_t E = ix.row_off[(size_t)n];
  if ((int64_t)ix.gdis.size() != E) bad("H2H labels of the wrong length");
  if ((int64_t)ix.sep_off.size() != (int64_t)n + 1 || ix.sep_off[0] != 0) bad("separator offsets");
  for (int32_t p = 0; p < n; ++p) {
    if (ix.sep_off[(size_t)p + 1] <= ix.sep_off[(size_t)p]) bad("a node without its own separator entry");
    for (int64_t i = ix.sep_off[(size_t)p]; i < ix.sep_off[(size_t)p + 1]; ++i)
      if ((int64_t)ix.sep.size() <= i || ix.sep[(size_t)i] < 0 || ix.sep[(size_t)i] > ix.depth[(size_t)p])
        bad("a separator depth out of range");
  }
  if ((int64_t)ix.sep.size() != ix.sep_off[(size_t)n]) bad("separator array of the wrong length");
  for (int64_t i = 0; i < E; ++i)
    if (ix.gdis[(size_t)i] < 0 || ix.gdis[(size_t)i] > kInf) bad("an H2H distance outside [0, INF]");
  if (ix.p2h_base.size() != ix.p2h_node.size()) bad("P2H tables of the wrong length");
  int64_t next = 0;
  for (size_t b = 0; b < ix.p2h_node.size(); ++b) {
    const int32_t v = ix.p2h_node[b];
    if (v < 0 || v >= n || (b > 0 && v <= ix.p2h_node[b - 1])) bad("P2H branch nodes out of order");
    if (ix.p2h_base[b] != next) bad("P2H table offsets");
    const int64_t size = ix.subtree[(size_t)v];
    next += size + 1;
    if ((int64_t)ix.p2h_off.size() < next) bad("P2H offsets of the wrong length");
    for (int64_t r = 0; r < size; ++r) {
      const int64_t o0 = ix.p2h_off[(size_t)(ix.p2h_base[b] + r)], o1 = ix.p2h_off[(size_t)(ix.p2h_base[b] + r + 1)];
      if (o0 < 0 || o1 < o0 || o1 > (int64_t)ix.p2h.size()) bad("P2H offsets are not monotone");
      for (int64_t i = o0; i < o1; ++i)
        if (ix.p2h[(size_t)i] < 0 || ix.p2h[(size_t)i] > ix.depth[(size_t)v]) bad("a P2H depth out of range");
    }
  }
  if ((int64_t)ix.p2h_off.size() != next) bad("P2H offsets of the wrong length");
}

// Checks an index assembled from outside (a loaded image or caller arrays)
// against every property the kernels and the host planning rely on, and
// derives iperm and subtree.  Throws BCX_EIO naming `what` on the first
// violation.
void validate_index(Index &ix, const std::string &what) {
  auto bad = [&](const std::string &m) { throw Error{BCX_EIO, what + ": " + m}; };
  const int32_t n = ix.n;
  if (n <= 0) bad("no vertices");
  if ((int64_t)ix.perm.size() != n || (int64_t)ix.depth.size() != n || (int64_t)ix.parent.size() != n)
    bad("tree arrays of the wrong length");
  ix.iperm.assign((size_t)n, -1);
  for (int32_t p = 0; p < n; ++p) {
    const int32_t v = ix.perm[(size_t)p];
    if (v < 0 || v >= n || ix.iperm[(size_t)v] != -1) bad("the preorder is not a permutation");
    ix.iperm[(size_t)v] = p;
  }
  if (ix.parent[0] != -1 || ix.depth[0] != 0) bad("node 0 is not the root");
  int32_t maxd = 0;
  for (int32_t p = 1; p < n; ++p) {
    const int32_t a = ix.parent[(size_t)p];
    if (a < 0 || a >= p) bad("a parent does not precede its child in preorder");
    if (ix.depth[(size_t)p] != ix.depth[(size_t)a] + 1) bad("a depth is not its parent's plus one");
    maxd = std::max(maxd, ix.depth[(size_t)p]);
  }
  if (ix.tree_height != maxd + 1) bad("tree height is not the largest depth plus one");
  ix.subtree.assign((size_t)n, 1);
  for (int32_t p = n - 1; p > 0; --p) ix.subtree[(size_t)ix.parent[(size_t)p]] += ix.subtree[(size_t)p];
  for (int32_t p = 1; p < n; ++p) {  // preorder: a subtree is a contiguous id range
    const int32_t a = ix.parent[(size_t)p];
    if (!(a < p && p + ix.subtree[(size_t)p] <= a + ix.subtree[(size_t)a])) bad("ids are not a preorder");
  }
  if ((int64_t)ix.row_off.size() != (int64_t)n + 1 || ix.row_off[0] != 0) bad("label offsets");
  for (int32_t p = 0; p < n; ++p)
    if (ix.row_off[(size_t)p + 1] - ix.row_off[(size_t)p] != (int64_t)ix.depth[(size_t)p] + 1)
      bad("a label row's length is not its depth plus one");
  const int64_t E = ix.row_off[(size_t)n];
  if ((int64_t)ix.dis.size() != E || (int64_t)ix.cnt.size() != E || (int64_t)ix.hop.size() != E)
    bad("label arrays of the wrong length");
  for (int32_t p = 0; p < n; ++p) {
    const int64_t o = ix.row_off[(size_t)p];
    const int32_t D = ix.depth[(size_t)p];
    if (ix.dis[(size_t)(o + D)] != 0 || ix.cnt[(size_t)(o + D)] != 1 || ix.hop[(size_t)(o + D)] != 0)
      bad("a label row's own entry is not (0, 1, 0)");
    for (int32_t j = 0; j < D; ++j) {
      const int32_t d = ix.dis[(size_t)(o + j)];
      if (d < 0 || d > kInf) bad("a label distance outside [0, 999999999]");
      if (d == kInf && ix.cnt[(size_t)(o + j)] != 0) bad("an unreachable label entry with a count");
    }
  }
  if ((int64_t)ix.bag_off.size() != (int64_t)n + 1 || ix.bag_off[0] != 0) bad("bag offsets");
  for (int32_t p = 0; p < n; ++p)
    if (ix.bag_off[(size_t)p + 1] < ix.bag_off[(size_t)p]) bad("bag offsets are not monotone");
  const int64_t B = ix.bag_off[(size_t)n];
  if ((int64_t)ix.bag.size() != B || (int64_t)ix.bag_len.size() != B || (int64_t)ix.bag_cnt.size() != B ||
      (int64_t)ix.bag_hop.size() != B)
    bad("bag arrays of the wrong length");
  int32_t width = 0;
  for (int32_t p = 0; p < n; ++p) {
    const int64_t e0 = ix.bag_off[(size_t)p], e1 = ix.bag_off[(size_t)p + 1];
    width = std::max(width, (int32_t)(e1 - e0));
    if (p == 0 ? e1 != e0 : e1 == e0) bad(p == 0 ? "the root has a bag" : "a non-root node has an empty bag");
    int32_t prev = -1;
    for (int64_t e = e0; e < e1; ++e) {
      const int32_t a = ix.bag[(size_t)e];
      if (a < 0 || a >= p || p >= a + ix.subtree[(size_t)a]) bad("a bag entry is not an ancestor");
      if (ix.depth[(size_t)a] <= prev) bad("bag entries are not in increasing depth");
      prev = ix.depth[(size_t)a];
      if (ix.bag_len[(size_t)e] < 1 || ix.bag_len[(size_t)e] > kInf) bad("a shortcut length outside [1, INF]");
      if (ix.bag_cnt[(size_t)e] == 0) bad("a shortcut with no path");
    }
    if (p > 0 && ix.bag[(size_t)(e1 - 1)] != ix.parent[(size_t)p]) bad("a node's deepest bag entry is not its parent");
  }
  if (ix.tree_width != width) bad("tree width is not the largest bag");
  if (!ix.sep_off.empty() || !ix.gdis.empty() || !ix.sep.empty()) check_h2h(ix, what);
  check_distance_bound(ix);
}

void save_index(const Index &ix, const char *path) {
  need_host_labels(ix, "saving an index image");
  FILE *f = std::fopen(path, "wb");
  if (!f) throw Error{BCX_EIO, std::string("cannot write ") + path};
  try {
    const int32_t hdr[4] = {ix.n, ix.tree_height, ix.tree_width, ix.count_overflow};
    const int64_t fp[2] = {ix.fp_arcs, (int64_t)ix.fp_hash};
    const int32_t h2h[2] = {ix.sep_off.empty() ? 0 : 1, 0};
    if (std::fwrite(kMagic, 1, 8, f) != 8 || std::fwrite(hdr, sizeof hdr, 1, f) != 1 ||
        std::fwrite(fp, sizeof fp, 1, f) != 1 || std::fwrite(h2h, sizeof h2h, 1, f) != 1)
      throw Error{BCX_EIO, "index write failed"};
    wvec(f, ix.perm);
    wvec(f, ix.depth);
    wvec(f, ix.parent);
    wvec(f, ix.row_off);
    wvec(f, ix.dis);
    wvec(f, ix.cnt);
    wvec(f, ix.hop);
    wvec(f, ix.bag_off);
    wvec(f, ix.bag);
    wvec(f, ix.bag_len);
    wvec(f, ix.bag_cnt);
    wvec(f, ix.bag_hop);
    if (h2h[0]) {
      wvec(f, ix.gdis);
      wvec(f, ix.sep_off);
      wvec(f, ix.sep);
      wvec(f, ix.p2h_node);
      wvec(f, ix.p2h_base);
      wvec(f, ix.p2h_off);
      wvec(f, ix.p2h);
    }
  } catch (...) {
    std::fclose(f);
    throw;
  }
  if (std::fclose(f) != 0) throw Error{BCX_EIO, "index write failed"};
}

void load_index(const char *path, Index &ix) {
  FILE *f = std::fopen(path, "rb");
  if (!f) throw Error{BCX_EIO, std::string("cannot open ") + path};
  ix = Index();
  try {
    char m[8];
    int32_t hdr[4], h2h[2];
    int64_t fp[2];
    if (std::fread(m, 1, 8, f) != 8 || std::memcmp(m, kMagic, 8) != 0 ||
        std::fread(hdr, sizeof hdr, 1, f) != 1 || std::fread(fp, sizeof fp, 1, f) != 1 ||
        std::fread(h2h, sizeof h2h, 1, f) != 1)
      throw Error{BCX_EIO, std::string(path) + " is not a bcx index image"};
    ix.n = hdr[0];
    ix.tree_height = hdr[1];
    ix.tree_width = hdr[2];
    ix.count_overflow = hdr[3];
    ix.fp_arcs = fp[0];
    ix.fp_hash = (uint64_t)fp[1];
    if (ix.n <= 0 || ix.tree_height <= 0 || ix.tree_height > ix.n)
      throw Error{BCX_EIO, std::string(path) + ": bad header"};
    const int64_t n = ix.n, cap = (int64_t)1 << 40;
    rvec(f, ix.perm, n);
    rvec(f, ix.depth, n);
    rvec(f, ix.parent, n);
    rvec(f, ix.row_off, n + 1);
    rvec(f, ix.dis, cap);
    rvec(f, ix.cnt, cap);
    rvec(f, ix.hop, cap);
    rvec(f, ix.bag_off, n + 1);
    rvec(f, ix.bag, cap);
    rvec(f, ix.bag_len, cap);
    rvec(f, ix.bag_cnt, cap);
    rvec(f, ix.bag_hop, cap);
    if (h2h[0]) {
      rvec(f, ix.gdis, cap);
      rvec(f, ix.sep_off, n + 1);
      rvec(f, ix.sep, cap);
      rvec(f, ix.p2h_node, n);
      rvec(f, ix.p2h_base, n);
      rvec(f, ix.p2h_off, cap);
      rvec(f, ix.p2h, cap);
    }
    char extra;
    if (std::fread(&extra, 1, 1, f) != 0) throw Error{BCX_EIO, std::string(path) + ": trailing data"};
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  validate_index(ix, path);
}

// ---- host arrays (the reference's TreeDecomp handed over) -----------------
// The tree, label rows and bags indexed by vertex id: parent[v] (-1 at the
// root); v's label row over its ancestors, root first (depth(v) entries, the
// reference's Node_::dis / cnt / my_cnt, TreeDecomp.h:95-118); v's bag, its
// neighbours at elimination with the shortcut's length, count and hop count
// (Node_::vertices, vertices_length, and TreeDecomp::edge_count_ /
// my_edge_count_, TreeDecomp.cpp:157-200).  Nodes are renumbered in DFS
// preorder with children in increasing vertex id, rows get their own entry
// (0, 1, 0) appended, bags are ordered by ancestor depth, and the result is
// validated like a loaded image.
void index_from_arrays(int32_t n, const int32_t *parent, const int64_t *label_off, const int32_t *dis,
                       const uint64_t *cnt, const uint32_t *hop, const int64_t *bag_off, const int32_t *bag_v,
                       const int32_t *bag_len, const uint64_t *bag_cnt, const uint32_t *bag_hop, Index &ix) {
  if (n <= 0) throw Error{BCX_EINVAL, "no vertices"};
  ix = Index();
  ix.n = n;
  int32_t root = -1;
  std::vector<int32_t> ch_off((size_t)n + 1, 0), ch((size_t)n);
  for (int32_t v = 0; v < n; ++v) {
    const int32_t a = parent[v];
    if (a == -1) {
      if (root >= 0) throw Error{BCX_EINVAL, "more than one root"};
      root = v;
    } else if (a < 0 || a >= n || a == v) {
      throw Error{BCX_EINVAL, "parent of vertex " + std::to_string(v) + " out of range"};
    } else {
      ch_off[(size_t)a + 1]++;
    }
  }
  if (root < 0) throw Error{BCX_EINVAL, "no root"};
  for (int32_t v = 0; v < n; ++v) ch_off[(size_t)v + 1] += ch_off[(size_t)v];
  {
    std::vector<int32_t> f(ch_off.begin(), ch_off.end() - 1);
    for (int32_t v = 0; v < n; ++v)  // increasing vertex id within each parent
      if (parent[v] >= 0) ch[(size_t)f[(size_t)parent[v]]++] = v;
  }
  ix.perm.assign((size_t)n, -1);
  ix.iperm.assign((size_t)n, -1);
  ix.depth.assign((size_t)n, 0);
  ix.parent.assign((size_t)n, -1);
  {
    std::vector<std::pair<int32_t, int32_t>> st{{root, -1}};  // (vertex, parent preorder id)
    int32_t next = 0;
    while (!st.empty()) {
      auto [v, pp] = st.back();
      st.pop_back();
      if (next >= n || ix.iperm[(size_t)v] != -1) throw Error{BCX_EINVAL, "parent array has a cycle"};
      const int32_t p = next++;
      ix.perm[(size_t)p] = v;
      ix.iperm[(size_t)v] = p;
      ix.parent[(size_t)p] = pp;
      ix.depth[(size_t)p] = pp < 0 ? 0 : ix.depth[(size_t)pp] + 1;
      for (int32_t k = ch_off[(size_t)v + 1]; k-- > ch_off[(size_t)v];) st.emplace_back(ch[(size_t)k], p);
    }
    if (next != n) throw Error{BCX_EINVAL, "parent array does not form one tree"};
  }
  ix.tree_height = 1 + *std::max_element(ix.depth.begin(), ix.depth.end());
  ix.row_off.assign((size_t)n + 1, 0);
  ix.bag_off.assign((size_t)n + 1, 0);
  for (int32_t p = 0; p < n; ++p) {
    const int32_t v = ix.perm[(size_t)p];
    if (label_off[v + 1] - label_off[v] != ix.depth[(size_t)p])
      throw Error{BCX_EINVAL, "label row of vertex " + std::to_string(v) + " does not have depth(v) entries"};
    if (bag_off[v + 1] < bag_off[v]) throw Error{BCX_EINVAL, "bag offsets are not monotone"};
    ix.row_off[(size_t)p + 1] = ix.row_off[(size_t)p] + ix.depth[(size_t)p] + 1;
    ix.bag_off[(size_t)p + 1] = ix.bag_off[(size_t)p] + (bag_off[v + 1] - bag_off[v]);
  }
  const size_t E = (size_t)ix.row_off[(size_t)n], B = (size_t)ix.bag_off[(size_t)n];
  ix.dis.resize(E);
  ix.cnt.resize(E);
  ix.hop.resize(E);
  ix.bag.resize(B);
  ix.bag_len.resize(B);
  ix.bag_cnt.resize(B);
  ix.bag_hop.resize(B);
  std::vector<int64_t> ord;
  for (int32_t p = 0; p < n; ++p) {
    const int32_t v = ix.perm[(size_t)p];
    const int64_t src = label_off[v], dst = ix.row_off[(size_t)p];
    const int32_t D = ix.depth[(size_t)p];
    for (int32_t j = 0; j < D; ++j) {
      ix.dis[(size_t)(dst + j)] = dis[src + j];
      ix.cnt[(size_t)(dst + j)] = cnt[src + j];
      ix.hop[(size_t)(dst + j)] = hop[src + j];
    }
    ix.dis[(size_t)(dst + D)] = 0;
    ix.cnt[(size_t)(dst + D)] = 1;
    ix.hop[(size_t)(dst + D)] = 0;
    ord.clear();
    for (int64_t e = bag_off[v]; e < bag_off[v + 1]; ++e) {
      if (bag_v[e] < 0 || bag_v[e] >= n) throw Error{BCX_EINVAL, "bag vertex out of range"};
      ord.push_back(e);
    }
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      return ix.depth[(size_t)ix.iperm[(size_t)bag_v[a]]] < ix.depth[(size_t)ix.iperm[(size_t)bag_v[b]]];
    });
    int64_t k = ix.bag_off[(size_t)p];
    for (int64_t e : ord) {
      ix.bag[(size_t)k] = ix.iperm[(size_t)bag_v[e]];
      ix.bag_len[(size_t)k] = bag_len[e];
      ix.bag_cnt[(size_t)k] = bag_cnt[e];
      ix.bag_hop[(size_t)k] = bag_hop[e];
      ++k;
    }
    ix.tree_width = std::max(ix.tree_width, (int32_t)ord.size());
  }
  for (size_t e = 0; e < E; ++e)
    if (ix.cnt[e] == UINT64_MAX) ix.count_overflow = 1;
  validate_index(ix, "index arrays");
}

void index_export(const Index &ix, int32_t *parent, int64_t *label_off, int32_t *dis, uint64_t *cnt,
                  uint32_t *hop, int64_t *bag_off, int32_t *bag_v, int32_t *bag_len, uint64_t *bag_cnt,
                  uint32_t *bag_hop) {
  if (dis || cnt || hop) need_host_labels(ix, "exporting label rows");
  const int32_t n = ix.n;
  int64_t lo = 0, bo = 0;
  for (int32_t v = 0; v < n; ++v) {
    const int32_t p = ix.iperm[(size_t)v];
    const int32_t D = ix.depth[(size_t)p];
    if (parent) parent[v] = ix.parent[(size_t)p] < 0 ? -1 : ix.perm[(size_t)ix.parent[(size_t)p]];
    if (label_off) label_off[v] = lo;
    const int64_t r = ix.row_off[(size_t)p];
    for (int32_t j = 0; j < D; ++j) {
      if (dis) dis[lo + j] = ix.dis[(size_t)(r + j)];
      if (cnt) cnt[lo + j] = ix.cnt[(size_t)(r + j)];
      if (hop) hop[lo + j] = ix.hop[(size_t)(r + j)];
    }
    lo += D;
    if (bag_off) bag_off[v] = bo;
    for (int64_t e = ix.bag_off[(size_t)p]; e < ix.bag_off[(size_t)p + 1]; ++e, ++bo) {
      if (bag_v) bag_v[bo] = ix.perm[(size_t)ix.bag[(size_t)e]];
      if (bag_len) bag_len[bo] = ix.bag_len[(size_t)e];
      if (bag_cnt) bag_cnt[bo] = ix.bag_cnt[(size_t)e];
      if (bag_hop) bag_hop[bo] = ix.bag_hop[(size_t)e];
    }
  }
  if (label_off) label_off[n] = lo;
  if (bag_off) bag_off[n] = bo;
}


// QBC's H2H index from the reference's host structure (Graph::Tree after
// makeIndex, optimizeVS and SVSP_P2H, buildIndex.cpp:208-278, 313-386,
// 565-719; the arrays moveDisIndexToDevice flattens, QBC.cu:11-127), indexed
// by vertex id, installed on ix's tree:
//   parent[v]          v's tree parent vertex (-1 at the root); must be ix's
//   gdis row of v      [gdis_off[v], gdis_off[v + 1]): depth(v) global
//                      distances to v's ancestors, root first (Node::dis);
//                      the own entry 0 is appended here
//   sep row of v       [sep_off[v], sep_off[v + 1]): row positions (= ancestor
//                      depths, v's own depth included) of the separator
//                      scanned when v is the LCA (Node::posOptimize), any order
//   branch[j]          the kept P2H branch vertices (Node::isBranchNode)
//   list of s below    [p2h_off[j * n + s], p2h_off[j * n + s + 1]) of p2h:
//   branch[j]          positions of branch[j]'s separator (Node::posP2H_3[s]);
//                      empty for s outside branch[j]'s subtree
// Positions must lie in the rows they index.  Every array is checked before
// anything of ix changes (BCX_EINVAL for shape, BCX_EIO for ranges).
void index_h2h_from_arrays(Index &ix, const int32_t *parent, const int64_t *gdis_off, const int32_t *gdis,
                           const int64_t *sep_off, const int32_t *sep, int32_t nbranch, const int32_t *branch,
                           const int64_t *p2h_off, const int32_t *p2h) {
  const int32_t n = ix.n;
  auto inval = [](const std::string &m) { throw Error{BCX_EINVAL, "H2H arrays: " + m}; };
  for (int32_t v = 0; v < n; ++v) {
    const int32_t pp = ix.parent[(size_t)ix.iperm[(size_t)v]];
    if (parent[v] != (pp < 0 ? -1 : ix.perm[(size_t)pp]))
      inval("the parent of vertex " + std::to_string(v) + " differs from the index's tree");
  }
  if (nbranch < 0 || nbranch > n) inval("branch count out of range");
  if (gdis_off[0] != 0 || sep_off[0] != 0 || (nbranch > 0 && p2h_off[0] != 0)) inval("offsets do not start at 0");
  Index h;  // the new H2H members, moved into ix once checked
  h.gdis.resize((size_t)ix.entries());
  h.sep_off.assign((size_t)n + 1, 0);
  for (int32_t p = 0; p < n; ++p) {
    const int32_t v = ix.perm[(size_t)p], D = ix.depth[(size_t)p];
    if (gdis_off[v + 1] - gdis_off[v] != D)
      inval("the distance row of vertex " + std::to_string(v) + " does not have depth(v) entries");
    const int64_t r = ix.row_off[(size_t)p];
    for (int32_t j = 0; j < D; ++j) h.gdis[(size_t)(r + j)] = gdis[gdis_off[v] + j];
    h.gdis[(size_t)(r + D)] = 0;
    if (sep_off[v + 1] < sep_off[v]) inval("separator offsets are not monotone");
    h.sep.insert(h.sep.end(), sep + sep_off[v], sep + sep_off[v + 1]);
    h.sep_off[(size_t)p + 1] = (int64_t)h.sep.size();
  }
  std::vector<std::pair<int32_t, int32_t>> br;  // (preorder id, j)
  for (int32_t j = 0; j < nbranch; ++j) {
    if (branch[j] < 0 || branch[j] >= n) inval("branch vertex out of range");
    br.emplace_back(ix.iperm[(size_t)branch[j]], j);
  }
  std::sort(br.begin(), br.end());
  for (size_t b = 1; b < br.size(); ++b)
    if (br[b].first == br[b - 1].first) inval("a branch vertex listed twice");
  for (auto [v, j] : br) {
    const int64_t *o = p2h_off + (int64_t)j * n;
    for (int32_t s = 0; s < n; ++s) {
      if (o[s + 1] < o[s]) inval("P2H offsets are not monotone");
      const int32_t q = ix.iperm[(size_t)s];
      if (o[s + 1] > o[s] && !(v <= q && q < v + ix.subtree[(size_t)v]))
        inval("a P2H list for a vertex outside its branch vertex's subtree");
    }
    h.p2h_node.push_back(v);
    h.p2h_base.push_back((int64_t)h.p2h_off.size());
    for (int32_t r = 0; r < ix.subtree[(size_t)v]; ++r) {
      const int32_t s = ix.perm[(size_t)(v + r)];
      h.p2h_off.push_back((int64_t)h.p2h.size());
      h.p2h.insert(h.p2h.end(), p2h + o[s], p2h + o[s + 1]);
    }
    h.p2h_off.push_back((int64_t)h.p2h.size());
  }
  auto swap_in = [&]() {
    std::swap(ix.gdis, h.gdis);
    std::swap(ix.sep_off, h.sep_off);
    std::swap(ix.sep, h.sep);
    std::swap(ix.p2h_node, h.p2h_node);
    std::swap(ix.p2h_base, h.p2h_base);
    std::swap(ix.p2h_off, h.p2h_off);
    std::swap(ix.p2h, h.p2h);
  };
  swap_in();
  try {
    check_h2h(ix, "H2H arrays");
  } catch (...) {
    swap_in();  // ix keeps what it had
    throw;
  }
}

void index_h2h_sizes(const Index &ix, int64_t &gdis_entries, int64_t &sep_entries, int32_t &nbranch,
                     int64_t &p2h_entries) {
  if (ix.sep_off.empty()) throw Error{BCX_EINVAL, "H2H labels not built"};
  gdis_entries = ix.entries() - ix.n;
  sep_entries = (int64_t)ix.sep.size();
  nbranch = (int32_t)ix.p2h_node.size();
  p2h_entries = (int64_t)ix.p2h.size();
}

void index_h2h_export(const Index &ix, int64_t *gdis_off, int32_t *gdis, int64_t *sep_off, int32_t *sep,
                      int32_t *branch, int64_t *p2h_off, int32_t *p2h) {
  if (ix.sep_off.empty()) throw Error{BCX_EINVAL, "H2H labels not built"};
  const int32_t n = ix.n;
  int64_t go = 0, so = 0;
  for (int32_t v = 0; v < n; ++v) {
    const int32_t p = ix.iperm[(size_t)v], D = ix.depth[(size_t)p];
    if (gdis_off) gdis_off[v] = go;
    if (gdis)
      for (int32_t j = 0; j < D; ++j) gdis[go + j] = ix.gdis[(size_t)(ix.row_off[(size_t)p] + j)];
    go += D;
    if (sep_off) sep_off[v] = so;
    for (int64_t i = ix.sep_off[(size_t)p]; i < ix.sep_off[(size_t)p + 1]; ++i, ++so)
      if (sep) sep[so] = ix.sep[(size_t)i];
  }
  if (gdis_off) gdis_off[n] = go;
  if (sep_off) sep_off[n] = so;
  int64_t po = 0;
  for (size_t b = 0; b < ix.p2h_node.size(); ++b) {
    const int32_t v = ix.p2h_node[b];
    if (branch) branch[b] = ix.perm[(size_t)v];
    for (int32_t s = 0; s < n; ++s) {
      const int32_t q = ix.iperm[(size_t)s];
      if (p2h_off) p2h_off[(int64_t)b * n + s] = po;
      if (v <= q && q < v + ix.subtree[(size_t)v]) {
        const int64_t i = ix.p2h_base[b] + (q - v);
        for (int64_t k = ix.p2h_off[(size_t)i]; k < ix.p2h_off[(size_t)i + 1]; ++k, ++po)
          if (p2h) p2h[po] = ix.p2h[(size_t)k];
      }
    }
  }
  if (p2h_off) p2h_off[(int64_t)ix.p2h_node.size() * n] = po;
}

}  // namespace bcx
