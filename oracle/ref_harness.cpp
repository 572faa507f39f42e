// ref_harness.cpp — dumps the reference's own host index for parity fixtures.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build_ref.sh together with the
// reference's host sources where they lie under the reference checkout
// (src/TreeCount/TreeDecomp.cpp, src/TreeCount/Graph.cpp, src/Graph.cpp,
// src/buildIndex.cpp); the binary goes to oracle/_ref/ and is never shipped.
//
//   tc_ref tc  <graph> [pairs]          TreeCount index (TreeDecomp.cpp:12-511)
//   tc_ref h2h <graph> <order> [pairs]  H2H index from an order file
//                                       (buildIndex.cpp:4-278)
//   tc_ref tcfull <graph>               TreeCount index, every node's label row
//                                       and bag (the data moveCountIndexToDevice
//                                       reads, MHLQBC.cu:6-128, plus the bags)
//   tc_ref h2hfull <graph> <order>      QBC's H2H index with its optimizeVS
//                                       separators and P2H branch tables (the
//                                       data moveDisIndexToDevice reads,
//                                       QBC.cu:11-127)
// Output (text): a stats line, the elimination order, one "V" line per vertex
// (id, parent vertex id, tree height, label length; 1-based ids), label rows
// of the vertices listed in the pairs file, and one "Q s t d count hop" line
// per pair.  tcfull prints, per vertex, "R v k d:c:h ..." (its label row over
// its k ancestors, root first) and "B v k u:len:c:h ..." (its bag: neighbour
// at elimination, shortcut length, count and hop count).  Pair queries evaluate the reference's scan (MHLQBC.cu:334-412 for
// TreeCount labels; min over the LCA's pos set for H2H distance labels) on the
// reference's own in-memory labels, with 64-bit count products.
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <set>
#include <string>
#include <utility>
#include <vector>
// The reference's include/Graph.h (Graph, TreeDecomp) is force-included by
// build_ref.sh so that no reference path is written into this file.

static std::vector<std::pair<int, int>> read_pairs(const char *path) {
  std::vector<std::pair<int, int>> p;
  if (!path) return p;
  std::ifstream f(path);
  int a, b;
  while (f >> a >> b) p.emplace_back(a, b);
  return p;
}

static int lca_rmq(const std::vector<std::vector<int>> &rmq, const int *epos,
                   const std::vector<int> &height, int x, int y) {
  int p = epos[x], q = epos[y];
  if (p > q) std::swap(p, q);
  int len = q - p + 1, i = 1, k = 0;
  while (i * 2 < len) {
    i *= 2;
    ++k;
  }
  q = q - i + 1;
  return height[rmq[k][p]] < height[rmq[k][q]] ? rmq[k][p] : rmq[k][q];
}

static int run_tc(const char *graph, const char *pairs_path) {
  TreeDecomp td(graph, "/dev/null");
  td.Reduce();
  td.MakeTree();
  td.MakeIndex();
  const int n = td.G_.n();
  long long entries = 0;
  for (auto &nd : td.tree_) entries += (long long)nd.dis.size();
  std::printf("n %d height %d width %d entries %lld\n", n, td.tree_height_, td.tree_width_, entries);
  std::printf("order");
  for (int v : td.delete_vertex_order_) std::printf(" %d", v);
  std::printf("\n");
  for (int v = 1; v <= n; ++v) {
    const auto &nd = td.tree_[td.rank_[v]];
    int pv = nd.parent < 0 ? 0 : td.tree_[nd.parent].vertex;
    std::printf("V %d %d %d %zu\n", v, pv, nd.height, nd.dis.size());
  }
  std::vector<int> height(td.tree_.size());
  for (size_t i = 0; i < td.tree_.size(); ++i) height[i] = td.tree_[i].height;
  auto pairs = read_pairs(pairs_path);
  std::set<int> rows;
  for (auto &pq : pairs) {
    rows.insert(pq.first);
    rows.insert(pq.second);
  }
  for (int v : rows) {
    const auto &nd = td.tree_[td.rank_[v]];
    std::printf("L %d %zu", v, nd.dis.size());
    for (size_t j = 0; j < nd.dis.size(); ++j)
      std::printf(" %d:%u:%u", nd.dis[j], (unsigned)nd.cnt[j], (unsigned)nd.my_cnt[j]);
    std::printf("\n");
  }
  for (auto &pq : pairs) {
    int p = pq.first, q = pq.second;
    if (p == q) {
      std::printf("Q %d %d 0 1 0\n", p, q);
      continue;
    }
    int x = td.rank_[p], y = td.rank_[q];
    int lca = lca_rmq(td.rmq_index_, td.euler_tour_pos_, height, x, y);
    int position = td.tree_[lca].pos.back();
    long long dis = 999999999;
    unsigned long long ret = 0;
    unsigned my_ret = 0;
    if (lca == x || lca == y) {
      if (lca == y) std::swap(x, y);
      int pos = td.tree_[x].pos.back();
      dis = td.tree_[y].dis[pos];
      ret = td.tree_[y].cnt[pos];
      my_ret = td.tree_[y].my_cnt[pos];
      --position;
    }
    long long tmp_dis = 999999999;
    unsigned long long cnt = 0;
    unsigned my_cnt = 0;
    const auto &X = td.tree_[x], &Y = td.tree_[y];
    for (int i = 0; i <= position; ++i) {
      long long t = (long long)X.dis[i] + Y.dis[i];
      unsigned long long c = (unsigned long long)X.cnt[i] * Y.cnt[i];
      unsigned h = X.my_cnt[i] + Y.my_cnt[i];
      if (tmp_dis > t) {
        tmp_dis = t;
        cnt = c;
        my_cnt = h;
      } else if (tmp_dis == t) {
        cnt += c;
        if (h > my_cnt) my_cnt = h;
      }
    }
    if (dis > tmp_dis) {
      dis = tmp_dis;
      ret = cnt;
      my_ret = my_cnt;
    } else if (dis == tmp_dis) {
      ret += cnt;
      if (my_cnt > my_ret) my_ret = my_cnt;
    }
    std::printf("Q %d %d %lld %llu %u\n", p, q, dis, ret, my_ret);
  }
  return 0;
}

static int run_tcfull(const char *graph) {
  TreeDecomp td(graph, "/dev/null");
  td.Reduce();
  td.MakeTree();
  td.MakeIndex();
  const int n = td.G_.n();
  std::printf("n %d height %d width %d\n", n, td.tree_height_, td.tree_width_);
  for (int v = 1; v <= n; ++v) {
    const auto &nd = td.tree_[td.rank_[v]];
    int pv = nd.parent < 0 ? 0 : td.tree_[nd.parent].vertex;
    std::printf("V %d %d %d %zu\n", v, pv, nd.height, nd.dis.size());
    std::printf("R %d %zu", v, nd.dis.size());
    for (size_t j = 0; j < nd.dis.size(); ++j)
      std::printf(" %d:%u:%u", nd.dis[j], (unsigned)nd.cnt[j], (unsigned)nd.my_cnt[j]);
    std::printf("\nB %d %zu", v, nd.vertices.size());
    for (size_t i = 0; i < nd.vertices.size(); ++i)
      std::printf(" %d:%d:%u:%u", nd.vertices[i], nd.vertices_length[i],
                  (unsigned)td.edge_count_[nd.vertex][i], (unsigned)td.my_edge_count_[nd.vertex][i]);
    std::printf("\n");
  }
  return 0;
}

static int run_h2h(const char *graph, const char *order, const char *pairs_path) {
  Graph g;
  g.ReadGraph(graph);
  g.CHconsorderMT(order);
  g.makeTree();
  g.makeIndex();
  const int n = g.nodenum;
  long long entries = 0;
  for (auto &nd : g.Tree) entries += (long long)nd.dis.size();
  std::printf("n %d height %d width %d entries %lld\n", n, g.heightMax, 0, entries);
  std::printf("order");
  for (int i = 0; i < n; ++i) std::printf(" %d", g.vNodeOrder[i] + 1);
  std::printf("\n");
  for (int v = 0; v < n; ++v) {
    const auto &nd = g.Tree[g.rank[v]];
    int pv = nd.pa < 0 ? 0 : g.Tree[nd.pa].uniqueVertex + 1;
    std::printf("V %d %d %d %zu\n", v + 1, pv, nd.height, nd.dis.size());
  }
  std::vector<int> height(g.Tree.size());
  for (size_t i = 0; i < g.Tree.size(); ++i) height[i] = g.Tree[i].height;
  auto pairs = read_pairs(pairs_path);
  for (auto &pq : pairs) {
    int p = pq.first - 1, q = pq.second - 1;
    if (p == q) {
      std::printf("Q %d %d 0 0 0\n", p + 1, q + 1);
      continue;
    }
    int x = g.rank[p], y = g.rank[q];
    int lca = lca_rmq(g.RMQIndex, g.toRMQ.data(), height, x, y);
    long long d;
    if (lca == x) {
      d = g.Tree[y].dis[g.Tree[x].pos.back()];
    } else if (lca == y) {
      d = g.Tree[x].dis[g.Tree[y].pos.back()];
    } else {
      d = 999999999LL * 4;
      for (int k : g.Tree[lca].pos) {
        long long t = (long long)g.Tree[x].dis[k] + g.Tree[y].dis[k];
        if (t < d) d = t;
      }
    }
    std::printf("Q %d %d %lld 0 0\n", p + 1, q + 1, d);
  }
  return 0;
}

// QBC's whole host H2H index, built as the QBC driver builds it before its
// upload (QBC.cu:421-434: order-file contraction, tree, global-distance labels,
// optimizeVS, SVSP_P2H with k = 10, d = 5, R = 1e6), i.e. the data
// moveDisIndexToDevice (QBC.cu:11-127) reads.  Per vertex (1-based ids):
//   "V v parent_vertex height row_length"
//   "D v k d_0 ... d_{k-1}"   Node::dis, root first (no own entry)
//   "S v k p_0 ... p_{k-1}"   Node::posOptimize, in the reference's order
// and per kept branch node b, per vertex s with a non-empty list:
//   "L b s k p_0 ... p_{k-1}" Node::posP2H_3[s]
static int run_h2hfull(const char *graph, const char *order) {
  Graph g;
  g.ReadGraph(graph);
  g.CHconsorderMT(order);
  g.makeTree();
  g.makeIndex();
  g.optimizeVS();
  std::vector<int> kept;
  g.SVSP_P2H(10, 5, 1000000, kept);
  const int n = g.nodenum;
  std::printf("n %d\n", n);
  for (int v = 0; v < n; ++v) {
    const auto &nd = g.Tree[g.rank[v]];
    int pv = nd.pa < 0 ? 0 : g.Tree[nd.pa].uniqueVertex + 1;
    std::printf("V %d %d %d %zu\nD %d %zu", v + 1, pv, nd.height, nd.dis.size(), v + 1, nd.dis.size());
    for (int d : nd.dis) std::printf(" %d", d);
    std::printf("\nS %d %zu", v + 1, nd.posOptimize.size());
    for (int p : nd.posOptimize) std::printf(" %d", p);
    std::printf("\n");
  }
  for (const auto &nd : g.Tree) {
    if (!nd.isBranchNode) continue;
    for (int s = 0; s < (int)nd.posP2H_3.size(); ++s) {
      const auto &l = nd.posP2H_3[s];
      if (l.empty()) continue;
      std::printf("L %d %d %zu", nd.uniqueVertex + 1, s + 1, l.size());
      for (int p : l) std::printf(" %d", p);
      std::printf("\n");
    }
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 3 && std::string(argv[1]) == "tc") return run_tc(argv[2], argc > 3 ? argv[3] : nullptr);
  if (argc >= 3 && std::string(argv[1]) == "tcfull") return run_tcfull(argv[2]);
  if (argc >= 4 && std::string(argv[1]) == "h2h")
    return run_h2h(argv[2], argv[3], argc > 4 ? argv[4] : nullptr);
  if (argc >= 4 && std::string(argv[1]) == "h2hfull") return run_h2hfull(argv[2], argv[3]);
  std::fprintf(stderr, "usage: tc_ref tc <graph> [pairs] | tc_ref tcfull <graph> | "
                       "tc_ref h2h <graph> <order> [pairs] | tc_ref h2hfull <graph> <order>\n");
  return 2;
}
