// Host-side exercise of the index builder, the graph reader and the CPU oracle,
// compiled by tests/test_host_sanitizers.py with AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only; no device code is built here).
//
//   host_sanitize_driver <graph file> <order file or -> <scratch dir>
//
// Builds both index kinds, checks their answers against the oracle's Dijkstra,
// round-trips the index image and both host-array seams, feeds them damaged
// input (every refusal must be an error, never a fault), and runs the oracle's
// Brandes, label-scan BC and RQ pair formula against each other.  Prints "ok"
// and exits 0, or names the first mismatch and exits 3.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "bcx.h"
#include "bcx_host.hpp"

extern "C" {
int64_t orc_sssp(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w, int32_t s, int64_t *d,
                 uint64_t *sigma, uint32_t *hop, int32_t *order, int32_t *overflow);
int orc_brandes(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w, const int32_t *sources,
                int32_t ns, double *bc, int32_t *overflow);
int orc_brandes_mt(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w, const int32_t *sources,
                   int32_t ns, double *bc, int32_t *overflow, int32_t threads);
int orc_label_bc(int32_t n, const int32_t *parent, const int64_t *label_off, const int32_t *dis,
                 const uint64_t *cnt, const uint32_t *hop, const int64_t *row, const int32_t *col, const int32_t *w,
                 const int32_t *sources, int32_t ns, double *bc, int32_t threads);
int orc_rq_pairs(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w, const int32_t *points,
                 int32_t q, double *out);
}

namespace {

[[noreturn]] void fail(const std::string &m) {
  std::printf("MISMATCH: %s\n", m.c_str());
  std::exit(3);
}

bool refused(void (*f)(void *), void *arg) {
  try {
    f(arg);
  } catch (const bcx::Error &) {
    return true;
  }
  return false;
}

// Sources spread over the vertex range (at most k).
std::vector<int32_t> spread(int32_t n, int32_t k) {
  std::vector<int32_t> s;
  for (int32_t i = 0; i < k && i < n; ++i) s.push_back((int32_t)((int64_t)i * n / std::min(k, n)));
  return s;
}

bool close(double a, double b) { return std::fabs(a - b) <= 1e-9 * std::max(1.0, std::fabs(b)); }

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: host_sanitize_driver <graph> <order|-> <scratch dir>\n");
    return 2;
  }
  const std::string dir = argv[3];
  bcx::Graph g;
  bcx::read_graph_file(argv[1], g);
  if (!bcx::graph_connected(g)) fail("graph not connected");
  const int32_t n = g.n;
  if (bcx::graph_fingerprint(g) == 0) fail("graph fingerprint 0");
  std::vector<int32_t> order;
  if (std::string(argv[2]) == "-")
    bcx::min_degree_order(g, order);
  else
    bcx::read_order_file(argv[2], n, order);

  // Oracle rows of a few sources.
  const std::vector<int32_t> src = spread(n, 24);
  std::vector<std::vector<int64_t>> D(src.size(), std::vector<int64_t>((size_t)n));
  std::vector<std::vector<uint64_t>> S(src.size(), std::vector<uint64_t>((size_t)n));
  std::vector<std::vector<uint32_t>> Hp(src.size(), std::vector<uint32_t>((size_t)n));
  for (size_t i = 0; i < src.size(); ++i) {
    int32_t ovf = 0;
    if (orc_sssp(n, g.row.data(), g.col.data(), g.w.data(), src[i], D[i].data(), S[i].data(), Hp[i].data(),
                 nullptr, &ovf) != n)
      fail("oracle did not settle every vertex");
  }

  // TreeCount index (min-degree policy) and the order-file index with H2H.
  bcx::Index mhl, ord;
  bcx::build_index(g, {}, mhl);
  bcx::build_index(g, order, ord);
  bcx::build_h2h(ord);
  for (size_t i = 0; i < src.size(); ++i)
    for (int32_t t = 0; t < n; t += 7) {
      int32_t d;
      uint64_t c;
      uint32_t h;
      bool ovf;
      bcx::index_query(mhl, src[i], t, d, c, h, ovf);
      if (d != D[i][(size_t)t] || c != S[i][(size_t)t] || h != Hp[i][(size_t)t])
        fail("MHL query " + std::to_string(src[i]) + " -> " + std::to_string(t));
      bcx::index_query(ord, src[i], t, d, c, h, ovf);
      if (d != D[i][(size_t)t] || c != S[i][(size_t)t]) fail("order-file MHL query");
      if (bcx::h2h_query(ord, src[i], t) != D[i][(size_t)t]) fail("H2H query");
    }
  double ms = 0, mp = 0;
  bcx::h2h_scan_stats(ord, 20000, 7, ms, mp);
  if (!(mp <= ms)) fail("P2H scans more than the separators");

  // Image round trip, then damaged images.
  const std::string img = dir + "/ix.bin";
  bcx::save_index(ord, img.c_str());
  bcx::Index back;
  bcx::load_index(img.c_str(), back);
  if (back.gdis != ord.gdis || back.dis != ord.dis || back.bag != ord.bag || back.p2h != ord.p2h)
    fail("image round trip");
  {
    FILE *f = std::fopen(img.c_str(), "rb");
    std::vector<unsigned char> bytes;
    for (int ch; (ch = std::fgetc(f)) != EOF;) bytes.push_back((unsigned char)ch);
    std::fclose(f);
    const std::string bad = dir + "/bad.bin";
    int refusals = 0;
    const int damaged = n > 20000 ? 6 : 48;  // images of large graphs are hundreds of MB
    for (int k = 0; k < damaged; ++k) {
      std::vector<unsigned char> b = bytes;
      const size_t at = (size_t)((k * 2654435761u) % b.size());
      if (k % 3 == 0)
        b.resize(at);  // truncated
      else
        b[at] ^= (unsigned char)(1u << (k % 8));  // one bit flipped
      FILE *o = std::fopen(bad.c_str(), "wb");
      std::fwrite(b.data(), 1, b.size(), o);
      std::fclose(o);
      bcx::Index x;
      try {
        bcx::load_index(bad.c_str(), x);
      } catch (const bcx::Error &) {
        ++refusals;
      }
    }
    if (refusals == 0) fail("no damaged image refused");
  }

  // Host-array seams: round trip, then damaged arrays.
  {
    const int64_t E = mhl.entries() - n, B = (int64_t)mhl.bag.size();
    std::vector<int32_t> par((size_t)n), dis((size_t)E), bv((size_t)B), bl((size_t)B);
    std::vector<int64_t> lo((size_t)n + 1), bo((size_t)n + 1);
    std::vector<uint64_t> cnt((size_t)E), bc((size_t)B);
    std::vector<uint32_t> hop((size_t)E), bh((size_t)B);
    bcx::index_export(mhl, par.data(), lo.data(), dis.data(), cnt.data(), hop.data(), bo.data(), bv.data(),
                      bl.data(), bc.data(), bh.data());
    bcx::Index re;
    bcx::index_from_arrays(n, par.data(), lo.data(), dis.data(), cnt.data(), hop.data(), bo.data(), bv.data(),
                           bl.data(), bc.data(), bh.data(), re);
    for (int32_t t = 0; t < n; t += 11) {
      int32_t d1, d2;
      uint64_t c1, c2;
      uint32_t h1, h2;
      bool o1, o2;
      bcx::index_query(mhl, src[0], t, d1, c1, h1, o1);
      bcx::index_query(re, src[0], t, d2, c2, h2, o2);
      if (d1 != d2 || c1 != c2 || h1 != h2) fail("MHL seam round trip");
    }
    struct Arg {
      int32_t n;
      std::vector<int32_t> par, dis, bv, bl;
      std::vector<int64_t> lo, bo;
      std::vector<uint64_t> cnt, bc;
      std::vector<uint32_t> hop, bh;
    };
    auto call = [](void *p) {
      auto *a = (Arg *)p;
      bcx::Index x;
      bcx::index_from_arrays(a->n, a->par.data(), a->lo.data(), a->dis.data(), a->cnt.data(), a->hop.data(),
                             a->bo.data(), a->bv.data(), a->bl.data(), a->bc.data(), a->bh.data(), x);
    };
    for (int k = 0; k < 4; ++k) {
      Arg a{n, par, dis, bv, bl, lo, bo, cnt, bc, hop, bh};
      if (k == 0) a.par[(size_t)(n / 2)] = n / 2;                      // a vertex its own parent
      if (k == 1) a.lo[1] += 1;                                       // a row one entry too long
      if (k == 2) a.bv[0] = n + 5;                                    // a bag vertex out of range
      if (k == 3) a.dis[0] = -4;                                      // a negative distance
      if (!refused(call, &a)) fail("damaged MHL arrays accepted (case " + std::to_string(k) + ")");
    }
    int64_t ge = 0, se = 0, pe = 0;
    int32_t nb = 0;
    bcx::index_h2h_sizes(ord, ge, se, nb, pe);
    std::vector<int64_t> go((size_t)n + 1), so((size_t)n + 1), po((size_t)nb * n + 1);
    std::vector<int32_t> gd((size_t)ge), sp((size_t)se), br((size_t)nb), pl((size_t)pe), opar((size_t)n);
    bcx::index_h2h_export(ord, go.data(), gd.data(), so.data(), sp.data(), br.data(), po.data(), pl.data());
    bcx::index_export(ord, opar.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr);
    bcx::Index o2;
    bcx::build_index(g, order, o2);
    bcx::index_h2h_from_arrays(o2, opar.data(), go.data(), gd.data(), so.data(), sp.data(), nb, br.data(),
                               po.data(), pl.data());
    if (o2.gdis != ord.gdis || o2.sep != ord.sep || o2.p2h != ord.p2h || o2.p2h_off != ord.p2h_off)
      fail("H2H seam round trip");
    if (nb > 0) {
      std::vector<int32_t> br2 = br;
      br2[0] = n + 1;
      try {
        bcx::index_h2h_from_arrays(o2, opar.data(), go.data(), gd.data(), so.data(), sp.data(), nb, br2.data(),
                                   po.data(), pl.data());
        fail("damaged H2H arrays accepted");
      } catch (const bcx::Error &) {
      }
      if (o2.gdis != ord.gdis || o2.p2h != ord.p2h) fail("a refused H2H hand-over changed the index");
    }
  }

  // Tree and bags without host rows.
  {
    bcx::Index dl;
    bcx::build_index(g, {}, dl, false);
    if (dl.bag != mhl.bag || dl.parent != mhl.parent) fail("device-label index tree or bags");
    try {
      bcx::need_host_labels(dl, "test");
      fail("an index without host rows claims them");
    } catch (const bcx::Error &) {
    }
  }

  // Oracle cross-checks: Brandes (one and several threads), the label-scan
  // method over the index's exported rows, the RQ pair formula.
  {
    const std::vector<int32_t> bs = spread(n, 40);
    std::vector<double> b1((size_t)n, 0.0), b2((size_t)n, 0.0), b3((size_t)n, 0.0);
    int32_t ovf = 0;
    if (orc_brandes(n, g.row.data(), g.col.data(), g.w.data(), bs.data(), (int32_t)bs.size(), b1.data(), &ovf))
      fail("orc_brandes");
    if (orc_brandes_mt(n, g.row.data(), g.col.data(), g.w.data(), bs.data(), (int32_t)bs.size(), b2.data(),
                       &ovf, 4))
      fail("orc_brandes_mt");
    const int64_t E = mhl.entries() - n;
    std::vector<int32_t> par((size_t)n), dis((size_t)E);
    std::vector<int64_t> lo((size_t)n + 1);
    std::vector<uint64_t> cnt((size_t)E);
    std::vector<uint32_t> hop((size_t)E);
    bcx::index_export(mhl, par.data(), lo.data(), dis.data(), cnt.data(), hop.data(), nullptr, nullptr, nullptr,
                      nullptr, nullptr);
    if (orc_label_bc(n, par.data(), lo.data(), dis.data(), cnt.data(), hop.data(), g.row.data(), g.col.data(),
                     g.w.data(), bs.data(), (int32_t)bs.size(), b3.data(), 3))
      fail("orc_label_bc");
    // Saturated path counts (the 35x35 unit grid) leave BC undefined: the
    // forms are then only run, not compared.
    for (int32_t v = 0; v < n && !ovf; ++v)
      if (!close(b2[(size_t)v], b1[(size_t)v]) || !close(b3[(size_t)v], b1[(size_t)v]))
        fail("oracle BC forms disagree at " + std::to_string(v));
    if (n <= 1500 && !ovf) {
      std::vector<int32_t> pts = spread(n, 6), all((size_t)n);
      for (int32_t v = 0; v < n; ++v) all[(size_t)v] = v;
      std::vector<double> rq(pts.size(), 0.0), full((size_t)n, 0.0);
      if (orc_rq_pairs(n, g.row.data(), g.col.data(), g.w.data(), pts.data(), (int32_t)pts.size(), rq.data()))
        fail("orc_rq_pairs");
      if (orc_brandes(n, g.row.data(), g.col.data(), g.w.data(), all.data(), n, full.data(), &ovf))
        fail("orc_brandes (all)");
      for (size_t i = 0; i < pts.size(); ++i)
        if (std::fabs(rq[i] - full[(size_t)pts[i]]) > 1e-6 * std::max(1.0, full[(size_t)pts[i]]))
          fail("RQ pairs vs Brandes");
    }
  }
  std::printf("ok n=%d height=%d width=%d\n", n, ord.tree_height, ord.tree_width);
  return 0;
}
