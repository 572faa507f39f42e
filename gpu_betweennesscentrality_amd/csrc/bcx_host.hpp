// Host-side data structures of the bcx library: the parsed graph and the
// tree-decomposition label index, both in the layout the device consumes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bcx {

// Reference INF for label distances (src/TreeCount/Util.h:35, TreeDecomp.h:58).
constexpr int32_t kInf = 999999999;

struct Graph {
  int32_t n = 0;
  // Directed arcs as listed in the file (0-based), grouped per tail in file
  // order: the adjacency Graph::ReadGraph builds (src/Graph.cpp:22-36).
  std::vector<int64_t> row;  // n + 1
  std::vector<int32_t> col;  // arcs
  std::vector<int32_t> w;    // arcs
  int64_t arcs() const { return (int64_t)col.size(); }
};

// Tree-decomposition label index.  Tree nodes are numbered in DFS preorder of
// the decomposition tree (children in creation order), so the subtree of node
// p is the contiguous id range [p, p + subtree_size[p]) and every ancestor of p
// has a smaller id.  Row p holds depth[p] + 1 entries: entry i < depth[p] is
// the label towards p's ancestor at depth i (root first), the last entry is a
// self entry (0, 1, 0) so that a pair query is one aggregation over the common
// prefix of two rows.
struct Index {
  int32_t n = 0;
  std::vector<int32_t> perm;    // preorder id -> vertex id
  std::vector<int32_t> iperm;   // vertex id -> preorder id
  std::vector<int32_t> depth;   // root = 0
  std::vector<int32_t> parent;  // preorder id of parent, -1 for the root
  std::vector<int32_t> subtree; // subtree sizes
  std::vector<int64_t> row_off; // n + 1 offsets into the label arrays
  std::vector<int32_t> dis;     // restricted (up-path) distance
  std::vector<uint64_t> cnt;    // number of shortest up-paths
  std::vector<uint32_t> hop;    // max hops among those paths
  // Bags: the upward shortcut neighbours of each node (all ancestors), ordered
  // by ancestor depth, with the shortcut's restricted length, number of
  // restricted shortest paths and max hop count at elimination time.
  std::vector<int64_t> bag_off;
  std::vector<int32_t> bag;       // preorder id of the ancestor
  std::vector<int32_t> bag_len;
  std::vector<uint64_t> bag_cnt;
  std::vector<uint32_t> bag_hop;
  // H2H distance labels of the QBC path, built on request (build_h2h): the
  // global shortest distance from each node to each ancestor and to itself
  // (rows as dis), and per node the separator depths scanned when the node is
  // the LCA of a pair whose ends lie in two different child subtrees.
  std::vector<int32_t> gdis;
  std::vector<int64_t> sep_off;  // n + 1 (empty until build_h2h)
  std::vector<int32_t> sep;      // ancestor depths, the node's own depth last
  // P2H branch tables (built with the H2H labels): for each kept branch node
  // b (preorder id p2h_node[b]) and each node s of its subtree, the depths of
  // b's separator that no other separator vertex dominates from s; list of s
  // = p2h[p2h_off[p2h_base[b] + s - p2h_node[b]] .. next).
  std::vector<int32_t> p2h_node;
  std::vector<int64_t> p2h_base;
  std::vector<int64_t> p2h_off;
  std::vector<int32_t> p2h;
  // Fingerprint of the graph the index was built for (arc count and an
  // order-independent hash of its arcs; 0 when unknown).
  int64_t fp_arcs = 0;
  uint64_t fp_hash = 0;
  int32_t tree_height = 0;  // max depth + 1
  int32_t tree_width = 0;   // max bag size
  int32_t count_overflow = 0;
  double build_seconds = 0.0;
  // Label rows are computed on the host (the default), or left to the device
  // (bcx_ctx_create computes them there; dis/cnt/hop then stay empty and the
  // host functions that read them refuse the index).
  bool host_labels = true;
  int64_t entries() const { return row_off.empty() ? 0 : row_off.back(); }
};

// Error carrier for the host side; the C ABI maps it to codes + message.
struct Error {
  int code;
  std::string msg;
};

// graph_io.cpp
void read_graph_file(const char *path, Graph &g);
void graph_from_arcs(int32_t n, int64_t arcs, const int32_t *u, const int32_t *v,
                     const int32_t *w, Graph &g);
bool graph_connected(const Graph &g);
// Order-independent hash of the arcs (u, v, w); never 0.
uint64_t graph_fingerprint(const Graph &g);
void read_order_file(const char *path, int32_t n, std::vector<int32_t> &order);

// index_build.cpp
// Elimination order: empty -> min-degree (reference Reduce policy); else
// order[v] = contraction rank of v (0 contracts first).
void build_index(const Graph &g, const std::vector<int32_t> &order, Index &ix,
                 bool host_labels = true);
// Throws BCX_EINVAL when the index's label rows are not on the host.
void need_host_labels(const Index &ix, const char *what);
void index_query(const Index &ix, int32_t s, int32_t t, int32_t &d,
                 uint64_t &sigma, uint32_t &hop, bool &ovf);
int32_t lca_depth(const Index &ix, int32_t p, int32_t q);  // preorder ids
// Min-degree contraction order of Graph::GraphConstract (buildIndex.cpp:
// 759-925): order[v] = rank; optionally the number of contracted-neighbour
// arcs (the reference's "super edge number") and the largest contracted
// neighbourhood seen before every 100,000th contraction (its progress lines).
void min_degree_order(const Graph &g, std::vector<int32_t> &order,
                      int64_t *shortcut_arcs = nullptr, std::vector<int32_t> *width_log = nullptr);
// H2H labels and separators (buildIndex.cpp:222-278, 313-386).
void build_h2h(Index &ix);
// Distance by the H2H separator scan (QBC.cu:271-331 without the P2H branch
// table); vertex ids; requires build_h2h.
int32_t h2h_query(const Index &ix, int32_t s, int32_t t);
// Mean scan length (label entries read) of the H2H distance query over
// `pairs` seeded random pairs, with and without the P2H tables.
void h2h_scan_stats(const Index &ix, int64_t pairs, uint64_t seed, double &mean_sep, double &mean_p2h);
// Tight pairs of every bag in `nodes` order (see index_build.cpp).
void tight_pairs(const Index &ix, const std::vector<int32_t> &nodes, std::vector<int32_t> &tz,
                 std::vector<int32_t> &tab, std::vector<double> &tfr);
void save_index(const Index &ix, const char *path);
void load_index(const char *path, Index &ix);
// Throws BCX_EGRAPH when shortest distances may reach kInf.
void check_distance_bound(const Index &ix);
// Validates an index from outside; derives iperm and subtree.
void validate_index(Index &ix, const std::string &what);
// Checks the H2H members of an index with a valid tree.
void check_h2h(const Index &ix, const std::string &what);
// The reference's host index structure, indexed by vertex id (see
// index_build.cpp), in and out.
void index_from_arrays(int32_t n, const int32_t *parent, const int64_t *label_off, const int32_t *dis,
                       const uint64_t *cnt, const uint32_t *hop, const int64_t *bag_off, const int32_t *bag_v,
                       const int32_t *bag_len, const uint64_t *bag_cnt, const uint32_t *bag_hop, Index &ix);
void index_export(const Index &ix, int32_t *parent, int64_t *label_off, int32_t *dis, uint64_t *cnt,
                  uint32_t *hop, int64_t *bag_off, int32_t *bag_v, int32_t *bag_len, uint64_t *bag_cnt,
                  uint32_t *bag_hop);
// QBC's H2H index (labels, separators, P2H tables) from the reference's host
// structure, indexed by vertex id, onto ix's tree; and out.
void index_h2h_from_arrays(Index &ix, const int32_t *parent, const int64_t *gdis_off, const int32_t *gdis,
                           const int64_t *sep_off, const int32_t *sep, int32_t nbranch, const int32_t *branch,
                           const int64_t *p2h_off, const int32_t *p2h);
void index_h2h_sizes(const Index &ix, int64_t &gdis_entries, int64_t &sep_entries, int32_t &nbranch,
                     int64_t &p2h_entries);
void index_h2h_export(const Index &ix, int64_t *gdis_off, int32_t *gdis, int64_t *sep_off, int32_t *sep,
                      int32_t *branch, int64_t *p2h_off, int32_t *p2h);

}  // namespace bcx
