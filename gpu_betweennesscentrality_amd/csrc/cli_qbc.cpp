// QBC <graph> <order> <deviceID> [outBC] [checkBC] — drop-in for the
// reference's QBC executable (QBC.cu:593-648): same arguments, stdout lines,
// BC file and check-file behaviour (exit 1 on |BC - expected| > 0.1).
// The index is contracted in the order file's order, as CHconsorderMT does
// (buildIndex.cpp:4-111); path counts and layers come from the count labels on
// the device instead of the reference's host Kahn pass (QBC.cu:485-555).
#include <cmath>

#include "cli_common.hpp"

int main(int argc, char *argv[]) {
  std::cout.precision(8);
  std::cout << std::fixed;
  if (argc < 4) {
    std::cerr << "usage: QBC <graph> <order> <deviceID> [outBC] [checkBC]" << std::endl;
    return 1;
  }
  const std::string road = argv[1], order = argv[2];
  const int device = cli::parse_device(argv[3]);
  bcx_graph *g = nullptr;
  cli::check(bcx_graph_read(road.c_str(), &g), "reading graph");
  int32_t n;
  int64_t m;
  bcx_graph_info(g, &n, &m);
  std::cout << "roadDataPath:" << road << "\t" << n << "\t" << m << "\n";

  auto t0 = std::chrono::high_resolution_clock::now();
  cli::Metrics mt{"QBC", road, n, m, device};
  std::vector<int32_t> range;
  const bool partial = cli::source_range(n, range);
  bcx_index *ix = nullptr;
  cli::check(bcx_index_build_flags(g, order.c_str(), nullptr, BCX_INDEX_DEVICE_LABELS, &ix),
             "building index from order file");
  mt.index_s = cli::seconds_since(t0);
  bcx_ctx *ctx = nullptr;
  cli::check(bcx_ctx_create(device, g, ix, &ctx), "creating device context");
  mt.context_s = cli::seconds_since(t0) - mt.index_s;
  mt.mode = cli::apply_env_tuning(ctx);
  std::vector<double> bc;
  uint32_t ovf = 0;
  mt.work = cli::compute_bc(ctx, mt.mode, n, partial ? &range : nullptr, bc, ovf);
  const double secs = cli::seconds_since(t0);
  mt.compute_s = secs - mt.index_s - mt.context_s;
  mt.total_s = secs;
  mt.overflow = ovf;
  std::cout << secs << "  seconds (---totalTime---)" << std::endl;
  if (ovf) std::cerr << "warning: a shortest-path count overflowed uint64" << std::endl;
  if (argc >= 5) cli::write_bc_file(argv[4], bc);
  cli::maybe_dump_bc(bc);
  cli::write_metrics(mt, ctx);
  if (argc >= 6) {
    std::ifstream ifs(argv[5]);
    if (!ifs) {
      std::cerr << "error: cannot open " << argv[5] << std::endl;
      return 1;
    }
    int cn, cm, vid;
    double v;
    ifs >> cn >> cm;
    std::vector<double> standard((size_t)std::max(cn, 0), 0.0);
    while (ifs >> vid >> v)
      if (vid >= 1 && vid <= cn) standard[(size_t)vid - 1] = v;
    for (int i = 0; i < n; ++i) {
      const double e = i < cn ? standard[(size_t)i] : 0.0;
      if (std::fabs(bc[(size_t)i] - e) > 1e-1) {
        std::cout << "BC value is not correct for node " << i << "\t" << bc[(size_t)i] << "\t" << e
                  << std::endl;
        return 1;
      }
    }
    std::cout << "BC value is correct\n\n" << std::endl;
  }
  bcx_ctx_free(ctx);
  bcx_index_free(ix);
  bcx_graph_free(g);
  return 0;
}
