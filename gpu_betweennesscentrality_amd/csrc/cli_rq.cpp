// RQ <graph> <queries> <deviceID> — drop-in for the reference's RQ executable
// (RQ.cu:542-610): same arguments, query-file contract (exit 1 on a short
// file) and stdout table of 1-based ids and BC values.
#include <iomanip>

#include "cli_common.hpp"

int main(int argc, char *argv[]) {
  std::cout.precision(8);
  std::cout << std::fixed;
  if (argc < 4) {
    std::cerr << "usage: RQ <graph> <queries> <deviceID>" << std::endl;
    return 1;
  }
  const std::string road = argv[1], qpath = argv[2];
  const int device = cli::parse_device(argv[3]);
  bcx_graph *g = nullptr;
  cli::check(bcx_graph_read(road.c_str(), &g), "reading graph");
  int32_t n;
  int64_t m;
  bcx_graph_info(g, &n, &m);
  std::cout << "roadName:" << road << "\t" << n << "\t" << m << "\n";

  // processBactch (RQ.cu:589-610)
  std::ifstream ifs(qpath);
  if (!ifs) {
    std::cout << "Cannot open: " << qpath << std::endl;
    return 1;
  }
  int q = 0;
  ifs >> q;
  std::vector<int32_t> pts;
  for (int i = 0; i < q; ++i) {
    int id;
    if (!(ifs >> id)) break;
    pts.push_back(id - 1);
  }
  if ((int)pts.size() != q) {
    std::cout << "randNumber.size()<" << pts.size() << "  querySize=" << q << std::endl;
    return 1;
  }
  std::cout << "The number of vertices to query: " << q << std::endl;
  for (int32_t p : pts)
    if (p < 0 || p >= n) {
      std::cerr << "error: query vertex " << p + 1 << " out of range" << std::endl;
      return 1;
    }
  if (q >= n) q = n;

  auto t0 = std::chrono::high_resolution_clock::now();
  std::cout << "Read Weighted Graph_ OK!\nn = " << n << ", m = " << m << std::endl;
  cli::Metrics mt{"RQ", road, n, m, device};
  mt.unit = "queries";
  bcx_index *ix = cli::mhl_index(g);
  mt.index_s = cli::seconds_since(t0);
  bcx_ctx *ctx = nullptr;
  cli::check(bcx_ctx_create(device, g, ix, &ctx), "creating device context");
  mt.context_s = cli::seconds_since(t0) - mt.index_s;
  mt.mode = BCX_MODE_SWEEP;  // bcx_subset_bc answers from a bag sweep
  cli::apply_env_tuning(ctx);
  std::vector<double> out((size_t)q);
  const int rc = bcx_subset_bc(ctx, pts.data(), q, out.data());  // 1: a count overflowed
  cli::check(rc, "computing subset BC");
  const double secs = cli::seconds_since(t0);
  mt.work = q;
  mt.overflow = rc == 1;
  mt.compute_s = secs - mt.index_s - mt.context_s;
  mt.total_s = secs;
  std::cout << secs << "  seconds (---totalTime---)" << std::endl;
  cli::write_metrics(mt, ctx);

  std::cout << std::left;
  std::cout << std::endl;
  std::cout << std::setw(10) << "vertexID" << std::setw(10) << "BCValue" << std::endl;
  std::cout << std::string(20, '-') << std::endl;
  for (int i = 0; i < q; ++i)
    std::cout << std::setw(10) << (pts[(size_t)i] + 1) << std::setw(10) << std::setw(10)
              << out[(size_t)i] << std::endl;
  std::cout << std::endl;
  bcx_ctx_free(ctx);
  bcx_index_free(ix);
  bcx_graph_free(g);
  return 0;
}
