// C ABI: graph and index entry points (host only).  Device entry points live
// in bcx_device.hip.  Each entry point converts exceptions into a return code
// and a thread-local message.
#include <cstring>
#include <new>
#include <random>
#include <string>

#include "bcx.h"
#include "bcx_host.hpp"
#include "bcx_internal.hpp"

struct bcx_graph {
  bcx::Graph g;
};
struct bcx_index {
  bcx::Index ix;
};

namespace bcx {
thread_local std::string g_last_error;
void set_error(const std::string &m) { g_last_error = m; }
const Graph &graph_of(const bcx_graph *g) { return g->g; }
const Index &index_of(const bcx_index *ix) { return ix->ix; }
}  // namespace bcx

extern "C" {

const char *bcx_last_error(void) { return bcx::g_last_error.c_str(); }
const char *bcx_version(void) { return "bcx 0.1 (gfx950)"; }

int bcx_graph_read(const char *path, bcx_graph **out) {
  BCX_GUARD_BEGIN
  if (!path || !out) return bcx::fail(BCX_EINVAL, "null argument");
  auto *h = new bcx_graph();
  try {
    bcx::read_graph_file(path, h->g);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_graph_from_arcs(int32_t n, int64_t arcs, const int32_t *u, const int32_t *v,
                        const int32_t *w, bcx_graph **out) {
  BCX_GUARD_BEGIN
  if (!out || (arcs > 0 && (!u || !v || !w))) return bcx::fail(BCX_EINVAL, "null argument");
  auto *h = new bcx_graph();
  try {
    bcx::graph_from_arcs(n, arcs, u, v, w, h->g);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_graph_info(const bcx_graph *g, int32_t *n, int64_t *arcs) {
  if (!g) return bcx::fail(BCX_EINVAL, "null graph");
  if (n) *n = g->g.n;
  if (arcs) *arcs = g->g.arcs();
  return BCX_OK;
}

int bcx_graph_connected(const bcx_graph *g) {
  if (!g) return bcx::fail(BCX_EINVAL, "null graph");
  return bcx::graph_connected(g->g) ? 1 : 0;
}

void bcx_graph_free(bcx_graph *g) { delete g; }

static int build_common(const bcx_graph *g, const std::vector<int32_t> &order, bcx_index **out,
                        uint32_t flags = 0) {
  if (!bcx::graph_connected(g->g))
    return bcx::fail(BCX_EGRAPH, "graph is disconnected; the index needs a connected graph");
  auto *h = new bcx_index();
  try {
    bcx::build_index(g->g, order, h->ix, (flags & BCX_INDEX_DEVICE_LABELS) == 0);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return BCX_OK;
}

int bcx_index_build_mhl(const bcx_graph *g, bcx_index **out) {
  BCX_GUARD_BEGIN
  if (!g || !out) return bcx::fail(BCX_EINVAL, "null argument");
  return build_common(g, {}, out);
  BCX_GUARD_END
}

int bcx_index_build_order(const bcx_graph *g, const char *order_path, bcx_index **out) {
  BCX_GUARD_BEGIN
  if (!g || !order_path || !out) return bcx::fail(BCX_EINVAL, "null argument");
  std::vector<int32_t> order;
  bcx::read_order_file(order_path, g->g.n, order);
  return build_common(g, order, out);
  BCX_GUARD_END
}

// order[v] = contraction rank of vertex v, checked to be a permutation.
static int copy_order(const bcx_graph *g, const int32_t *order, std::vector<int32_t> &ord) {
  const int32_t n = g->g.n;
  ord.assign(order, order + n);
  std::vector<char> used((size_t)n, 0);
  for (int32_t v = 0; v < n; ++v) {
    if (ord[(size_t)v] < 0 || ord[(size_t)v] >= n || used[(size_t)ord[(size_t)v]])
      return bcx::fail(BCX_EINVAL, "order is not a permutation of 0..n-1");
    used[(size_t)ord[(size_t)v]] = 1;
  }
  return BCX_OK;
}

int bcx_index_build_flags(const bcx_graph *g, const char *order_path, const int32_t *order, uint32_t flags,
                          bcx_index **out) {
  BCX_GUARD_BEGIN
  if (!g || !out) return bcx::fail(BCX_EINVAL, "null argument");
  if (order_path && order) return bcx::fail(BCX_EINVAL, "give an order file or an order array, not both");
  if (flags & ~(uint32_t)BCX_INDEX_DEVICE_LABELS) return bcx::fail(BCX_EINVAL, "unknown index flags");
  std::vector<int32_t> ord;
  if (order_path) bcx::read_order_file(order_path, g->g.n, ord);
  if (order && copy_order(g, order, ord) != BCX_OK) return BCX_EINVAL;
  return build_common(g, ord, out, flags);
  BCX_GUARD_END
}

int bcx_index_build_with_order(const bcx_graph *g, const int32_t *order, bcx_index **out) {
  BCX_GUARD_BEGIN
  if (!g || !order || !out) return bcx::fail(BCX_EINVAL, "null argument");
  std::vector<int32_t> ord;
  if (copy_order(g, order, ord) != BCX_OK) return BCX_EINVAL;
  return build_common(g, ord, out);
  BCX_GUARD_END
}

int bcx_min_degree_order(const bcx_graph *g, int32_t *order, int64_t *shortcut_arcs,
                         int32_t *width_log) {
  BCX_GUARD_BEGIN
  if (!g || !order) return bcx::fail(BCX_EINVAL, "null argument");
  std::vector<int32_t> o, wl;
  bcx::min_degree_order(g->g, o, shortcut_arcs, width_log ? &wl : nullptr);
  std::memcpy(order, o.data(), o.size() * sizeof(int32_t));
  if (width_log) std::memcpy(width_log, wl.data(), wl.size() * sizeof(int32_t));
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_build_h2h(bcx_index *ix) {
  BCX_GUARD_BEGIN
  if (!ix) return bcx::fail(BCX_EINVAL, "null index");
  bcx::build_h2h(ix->ix);
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_h2h_query(const bcx_index *ix, int32_t s, int32_t t, int32_t *d) {
  BCX_GUARD_BEGIN
  if (!ix || !d) return bcx::fail(BCX_EINVAL, "null argument");
  const auto &x = ix->ix;
  if (s < 0 || s >= x.n || t < 0 || t >= x.n) return bcx::fail(BCX_EINVAL, "vertex out of range");
  if (x.sep_off.empty()) return bcx::fail(BCX_EINVAL, "H2H labels not built (bcx_index_build_h2h)");
  *d = bcx::h2h_query(x, s, t);
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_h2h_scan_stats(const bcx_index *ix, int64_t pairs, uint64_t seed, double *mean_separator,
                             double *mean_p2h) {
  BCX_GUARD_BEGIN
  if (!ix || pairs < 0) return bcx::fail(BCX_EINVAL, "bad argument");
  double a = 0, b = 0;
  bcx::h2h_scan_stats(ix->ix, pairs, seed, a, b);
  if (mean_separator) *mean_separator = a;
  if (mean_p2h) *mean_p2h = b;
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_stats_get(const bcx_index *ix, bcx_index_stats *st) {
  if (!ix || !st) return bcx::fail(BCX_EINVAL, "null argument");
  const auto &x = ix->ix;
  st->n = x.n;
  st->tree_height = x.tree_height;
  st->tree_width = x.tree_width;
  st->count_overflow = x.count_overflow;
  st->label_entries = x.entries();
  st->mean_depth = x.n ? (double)x.entries() / x.n : 0.0;
  st->build_seconds = x.build_seconds;
  st->bag_entries = x.bag_off.empty() ? 0 : x.bag_off.back();
  return BCX_OK;
}

int bcx_index_query(const bcx_index *ix, int32_t s, int32_t t, int32_t *d, uint64_t *sigma,
                    uint32_t *hop) {
  if (!ix) return bcx::fail(BCX_EINVAL, "null index");
  const auto &x = ix->ix;
  if (!x.host_labels)
    return bcx::fail(BCX_EINVAL, "index query needs the label rows on the host; this index leaves them to the device");
  if (s < 0 || s >= x.n || t < 0 || t >= x.n) return bcx::fail(BCX_EINVAL, "vertex out of range");
  int32_t dd;
  uint64_t ss;
  uint32_t hh;
  bool ovf;
  bcx::index_query(x, s, t, dd, ss, hh, ovf);
  if (d) *d = dd;
  if (sigma) *sigma = ss;
  if (hop) *hop = hh;
  return ovf ? 1 : BCX_OK;
}

int bcx_index_tree_order(const bcx_index *ix, int32_t *order) {
  if (!ix || !order) return bcx::fail(BCX_EINVAL, "null argument");
  std::memcpy(order, ix->ix.perm.data(), ix->ix.perm.size() * sizeof(int32_t));
  return BCX_OK;
}

int bcx_index_tree(const bcx_index *ix, int32_t *parent, int32_t *height) {
  if (!ix || !parent || !height) return bcx::fail(BCX_EINVAL, "null argument");
  const auto &x = ix->ix;
  for (int32_t p = 0; p < x.n; ++p) {
    const int32_t v = x.perm[(size_t)p];
    parent[v] = x.parent[(size_t)p] < 0 ? -1 : x.perm[(size_t)x.parent[(size_t)p]];
    height[v] = x.depth[(size_t)p] + 1;
  }
  return BCX_OK;
}

int bcx_index_mean_scan(const bcx_index *ix, int64_t pairs, uint64_t seed, double *mean) {
  if (!ix || !mean || pairs <= 0) return bcx::fail(BCX_EINVAL, "bad argument");
  const auto &x = ix->ix;
  std::mt19937_64 rng(seed);
  double acc = 0;
  for (int64_t k = 0; k < pairs; ++k) {
    int32_t p = (int32_t)(rng() % (uint64_t)x.n), q = (int32_t)(rng() % (uint64_t)x.n);
    acc += bcx::lca_depth(x, std::min(p, q), std::max(p, q)) + 1;
  }
  *mean = acc / (double)pairs;
  return BCX_OK;
}

int bcx_index_save(const bcx_index *ix, const char *path) {
  BCX_GUARD_BEGIN
  if (!ix || !path) return bcx::fail(BCX_EINVAL, "null argument");
  bcx::save_index(ix->ix, path);
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_load(const char *path, bcx_index **out) {
  BCX_GUARD_BEGIN
  if (!path || !out) return bcx::fail(BCX_EINVAL, "null argument");
  auto *h = new bcx_index();
  try {
    bcx::load_index(path, h->ix);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_from_arrays(int32_t n, const int32_t *parent, const int64_t *label_off, const int32_t *dis,
                          const uint64_t *cnt, const uint32_t *hop, const int64_t *bag_off, const int32_t *bag_v,
                          const int32_t *bag_len, const uint64_t *bag_cnt, const uint32_t *bag_hop,
                          const bcx_graph *g, bcx_index **out) {
  BCX_GUARD_BEGIN
  if (!parent || !label_off || !dis || !cnt || !hop || !bag_off || !bag_v || !bag_len || !bag_cnt || !bag_hop ||
      !out)
    return bcx::fail(BCX_EINVAL, "null argument");
  if (g && g->g.n != n) return bcx::fail(BCX_EINVAL, "graph and arrays have different vertex counts");
  auto *h = new bcx_index();
  try {
    bcx::index_from_arrays(n, parent, label_off, dis, cnt, hop, bag_off, bag_v, bag_len, bag_cnt, bag_hop, h->ix);
    if (g) {
      h->ix.fp_arcs = g->g.arcs();
      h->ix.fp_hash = bcx::graph_fingerprint(g->g);
    }
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_export(const bcx_index *ix, int32_t *parent, int64_t *label_off, int32_t *dis, uint64_t *cnt,
                     uint32_t *hop, int64_t *bag_off, int32_t *bag_v, int32_t *bag_len, uint64_t *bag_cnt,
                     uint32_t *bag_hop) {
  BCX_GUARD_BEGIN
  if (!ix) return bcx::fail(BCX_EINVAL, "null index");
  bcx::index_export(ix->ix, parent, label_off, dis, cnt, hop, bag_off, bag_v, bag_len, bag_cnt, bag_hop);
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_h2h_from_arrays(bcx_index *ix, const int32_t *parent, const int64_t *gdis_off,
                              const int32_t *gdis, const int64_t *sep_off, const int32_t *sep,
                              int32_t nbranch, const int32_t *branch, const int64_t *p2h_off,
                              const int32_t *p2h) {
  BCX_GUARD_BEGIN
  if (!ix || !parent || !gdis_off || !gdis || !sep_off || !sep ||
      (nbranch > 0 && (!branch || !p2h_off || !p2h)))
    return bcx::fail(BCX_EINVAL, "null argument");
  bcx::index_h2h_from_arrays(ix->ix, parent, gdis_off, gdis, sep_off, sep, nbranch, branch, p2h_off, p2h);
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_h2h_sizes(const bcx_index *ix, int64_t *gdis_entries, int64_t *sep_entries, int32_t *nbranch,
                        int64_t *p2h_entries) {
  BCX_GUARD_BEGIN
  if (!ix) return bcx::fail(BCX_EINVAL, "null index");
  int64_t a = 0, b = 0, d = 0;
  int32_t c = 0;
  bcx::index_h2h_sizes(ix->ix, a, b, c, d);
  if (gdis_entries) *gdis_entries = a;
  if (sep_entries) *sep_entries = b;
  if (nbranch) *nbranch = c;
  if (p2h_entries) *p2h_entries = d;
  return BCX_OK;
  BCX_GUARD_END
}

int bcx_index_h2h_export(const bcx_index *ix, int64_t *gdis_off, int32_t *gdis, int64_t *sep_off, int32_t *sep,
                         int32_t *branch, int64_t *p2h_off, int32_t *p2h) {
  BCX_GUARD_BEGIN
  if (!ix) return bcx::fail(BCX_EINVAL, "null index");
  bcx::index_h2h_export(ix->ix, gdis_off, gdis, sep_off, sep, branch, p2h_off, p2h);
  return BCX_OK;
  BCX_GUARD_END
}

void bcx_index_free(bcx_index *ix) { delete ix; }

}  // extern "C"
