// Graph, order and query file readers for the bcx library.
//
// Graph file contract (README.md:17-19, src/Graph.cpp:11-36): "n m", then m
// lines "u v w" with 1-based ids.  The reference builds two views of the file:
// a directed adjacency in file order with first-wins duplicate removal and
// w = 0 -> 1 (src/Graph.cpp:22-36) and an undirected min-weight map
// (src/TreeCount/Graph.cpp:55-71).  The two agree only when every edge is
// listed both ways with one weight >= 1, so that is enforced here: anything
// else is reported as BCX_EGRAPH instead of producing inconsistent indexes.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "bcx_host.hpp"
#include "bcx.h"

namespace bcx {

namespace {

struct Reader {
  std::vector<char> buf;
  size_t pos = 0;
  bool next_int(int64_t &out) {
    while (pos < buf.size() && (buf[pos] == ' ' || buf[pos] == '\n' ||
                                buf[pos] == '\r' || buf[pos] == '\t'))
      ++pos;
    if (pos >= buf.size()) return false;
    bool neg = false;
    if (buf[pos] == '-' || buf[pos] == '+') {
      neg = buf[pos] == '-';
      ++pos;
    }
    if (pos >= buf.size() || buf[pos] < '0' || buf[pos] > '9') return false;
    int64_t v = 0;
    while (pos < buf.size() && buf[pos] >= '0' && buf[pos] <= '9') {
      v = v * 10 + (buf[pos] - '0');
      if (v > (int64_t)1 << 40) return false;
      ++pos;
    }
    out = neg ? -v : v;
    return true;
  }
};

void slurp(const char *path, Reader &r) {
  FILE *f = std::fopen(path, "rb");
  if (!f) throw Error{BCX_EIO, std::string("cannot open ") + path + ": " + std::strerror(errno)};
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    throw Error{BCX_EIO, std::string("cannot size ") + path};
  }
  r.buf.resize((size_t)sz);
  size_t got = sz ? std::fread(r.buf.data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  if (got != (size_t)sz) throw Error{BCX_EIO, std::string("short read on ") + path};
}

}  // namespace

void graph_from_arcs(int32_t n, int64_t m, const int32_t *u, const int32_t *v,
                     const int32_t *w, Graph &g) {
  if (n <= 0) throw Error{BCX_EGRAPH, "graph has no vertices"};
  if (m < 0) throw Error{BCX_EINVAL, "negative arc count"};
  for (int64_t i = 0; i < m; ++i) {
    if (u[i] < 0 || u[i] >= n || v[i] < 0 || v[i] >= n)
      throw Error{BCX_EGRAPH, "arc " + std::to_string(i) + " has a vertex id out of range"};
    if (u[i] == v[i])
      throw Error{BCX_EGRAPH, "arc " + std::to_string(i) + " is a self loop"};
    if (w[i] <= 0)
      throw Error{BCX_EGRAPH, "arc " + std::to_string(i) + " has weight <= 0"};
  }
  // Arc ids bucketed by tail in file order (a counting sort), then, per tail,
  // ordered by (head, file position) to find duplicates and reverses: the same
  // (u, v, file position) order as one global sort, at linear cost.
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int64_t i = 0; i < m; ++i) off[(size_t)u[i] + 1]++;
  for (int32_t x = 0; x < n; ++x) off[(size_t)x + 1] += off[(size_t)x];
  std::vector<int64_t> ord((size_t)m);
  {
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t i = 0; i < m; ++i) ord[(size_t)fill[(size_t)u[i]]++] = i;
  }
  for (int32_t x = 0; x < n; ++x)
    std::sort(ord.begin() + off[(size_t)x], ord.begin() + off[(size_t)x + 1], [&](int64_t a, int64_t b) {
      return v[a] != v[b] ? v[a] < v[b] : a < b;
    });
  std::vector<char> keep((size_t)m, 1);
  for (size_t k = 1; k < ord.size(); ++k) {
    int64_t a = ord[k - 1], b = ord[k];
    if (u[a] == u[b] && v[a] == v[b]) {
      if (w[a] != w[b])
        throw Error{BCX_EGRAPH, "arc " + std::to_string(u[b] + 1) + " " + std::to_string(v[b] + 1) +
                                    " listed twice with different weights"};
      keep[(size_t)b] = 0;  // first listing wins
    }
  }
  // Every kept arc needs its reverse with the same weight: the first listing
  // of (v, u) in v's bucket.
  for (int64_t a : ord) {
    if (!keep[(size_t)a]) continue;
    const int32_t ru = v[a], rv = u[a];
    const auto lo = ord.begin() + off[(size_t)ru], hi = ord.begin() + off[(size_t)ru + 1];
    auto it = std::lower_bound(lo, hi, rv, [&](int64_t x, int32_t key) { return v[x] < key; });
    if (it == hi || v[*it] != rv)
      throw Error{BCX_EGRAPH, "arc " + std::to_string(u[a] + 1) + " " + std::to_string(v[a] + 1) +
                                  " has no reverse arc (edges must be listed both ways)"};
    if (w[*it] != w[a])
      throw Error{BCX_EGRAPH, "edge " + std::to_string(u[a] + 1) + " " + std::to_string(v[a] + 1) +
                                  " has different weights in its two directions"};
  }
  // Shortest distances must stay below the label INF (999999999); the index
  // build checks that bound on the distances themselves.
  g.n = n;
  g.row.assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < m; ++i)
    if (keep[(size_t)i]) g.row[(size_t)u[i] + 1]++;
  for (int32_t i = 0; i < n; ++i) g.row[(size_t)i + 1] += g.row[(size_t)i];
  g.col.assign((size_t)g.row[(size_t)n], 0);
  g.w.assign((size_t)g.row[(size_t)n], 0);
  std::vector<int64_t> fill(g.row.begin(), g.row.end() - 1);
  for (int64_t i = 0; i < m; ++i) {  // file order within each tail
    if (!keep[(size_t)i]) continue;
    int64_t p = fill[(size_t)u[i]]++;
    g.col[(size_t)p] = v[i];
    g.w[(size_t)p] = w[i];
  }
}

void read_graph_file(const char *path, Graph &g) {
  Reader r;
  slurp(path, r);
  int64_t n, m;
  if (!r.next_int(n) || !r.next_int(m) || n <= 0 || n > 0x7fffffff || m < 0)
    throw Error{BCX_EINVAL, std::string(path) + ": bad header (expected \"n m\")"};
  std::vector<int32_t> u((size_t)m), v((size_t)m), w((size_t)m);
  for (int64_t i = 0; i < m; ++i) {
    int64_t a, b, c;
    if (!r.next_int(a) || !r.next_int(b) || !r.next_int(c))
      throw Error{BCX_EINVAL, std::string(path) + ": truncated at arc line " + std::to_string(i + 1)};
    if (a < 1 || a > n || b < 1 || b > n)
      throw Error{BCX_EGRAPH, std::string(path) + ": vertex id out of range at arc line " +
                                  std::to_string(i + 1)};
    if (c < -0x7fffffffLL || c > 0x7fffffffLL)
      throw Error{BCX_EGRAPH, std::string(path) + ": weight out of int32 range"};
    u[(size_t)i] = (int32_t)(a - 1);
    v[(size_t)i] = (int32_t)(b - 1);
    w[(size_t)i] = (int32_t)c;
  }
  graph_from_arcs((int32_t)n, m, u.data(), v.data(), w.data(), g);
}

uint64_t graph_fingerprint(const Graph &g) {
  uint64_t h = 0;
  for (int32_t x = 0; x < g.n; ++x)
    for (int64_t k = g.row[(size_t)x]; k < g.row[(size_t)x + 1]; ++k) {
      uint64_t z = ((uint64_t)(uint32_t)x << 32 | (uint32_t)g.col[(size_t)k]) ^
                   ((uint64_t)(uint32_t)g.w[(size_t)k] * 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;  // splitmix64 finaliser
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      h += z ^ (z >> 31);
    }
  return h ? h : 1;
}

bool graph_connected(const Graph &g) {
  if (g.n <= 1) return true;
  std::vector<char> seen((size_t)g.n, 0);
  std::vector<int32_t> st{0};
  seen[0] = 1;
  int32_t cnt = 1;
  while (!st.empty()) {
    int32_t x = st.back();
    st.pop_back();
    for (int64_t k = g.row[(size_t)x]; k < g.row[(size_t)x + 1]; ++k) {
      int32_t y = g.col[(size_t)k];
      if (!seen[(size_t)y]) {
        seen[(size_t)y] = 1;
        ++cnt;
        st.push_back(y);
      }
    }
  }
  return cnt == g.n;
}

// Order file (README.md:22-23, buildIndex.cpp:11-20): "num" then num lines
// "nodeID order", 0-based ids; contraction runs in increasing order.
void read_order_file(const char *path, int32_t n, std::vector<int32_t> &order) {
  Reader r;
  slurp(path, r);
  int64_t num;
  if (!r.next_int(num) || num < 0)
    throw Error{BCX_EINVAL, std::string(path) + ": bad header (expected vertex count)"};
  order.assign((size_t)n, -1);
  std::vector<char> used((size_t)n, 0);
  for (int64_t i = 0; i < num; ++i) {
    int64_t id, ord;
    if (!r.next_int(id) || !r.next_int(ord))
      throw Error{BCX_EINVAL, std::string(path) + ": truncated at line " + std::to_string(i + 2)};
    if (id < 0 || id >= n || ord < 0 || ord >= n)
      throw Error{BCX_EINVAL, std::string(path) + ": id or order out of range at line " +
                                  std::to_string(i + 2)};
    if (used[(size_t)ord] || order[(size_t)id] != -1)
      throw Error{BCX_EINVAL, std::string(path) + ": repeated vertex or order value at line " +
                                  std::to_string(i + 2)};
    used[(size_t)ord] = 1;
    order[(size_t)id] = (int32_t)ord;
  }
  for (int32_t v = 0; v < n; ++v)
    if (order[(size_t)v] < 0)
      throw Error{BCX_EINVAL, std::string(path) + ": vertex " + std::to_string(v) + " has no order"};
}

}  // namespace bcx
