// MHLQBC <graph> <deviceID> [outBC] — drop-in for the reference's MHL-QBC
// executable (MHLQBC.cu:547-601): same arguments, stdout lines and BC file.
#include "cli_common.hpp"

int main(int argc, char *argv[]) {
  std::cout.precision(8);
  std::cout << std::fixed;
  if (argc < 3) {
    std::cerr << "usage: MHLQBC <graph> <deviceID> [outBC]" << std::endl;
    return 1;
  }
  const std::string road = argv[1];
  const int device = cli::parse_device(argv[2]);
  bcx_graph *g = nullptr;
  cli::check(bcx_graph_read(road.c_str(), &g), "reading graph");
  int32_t n;
  int64_t m;
  bcx_graph_info(g, &n, &m);
  std::cout << "roadDataPath:" << road << "\t" << n << "\t" << m << "\n";

  auto t0 = std::chrono::high_resolution_clock::now();
  std::cout << "Read Weighted Graph_ OK!\nn = " << n << ", m = " << m << std::endl;
  cli::Metrics mt{"MHLQBC", road, n, m, device};
  std::vector<int32_t> range;
  const bool partial = cli::source_range(n, range);
  bcx_index *ix = cli::mhl_index(g);
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
  if (argc >= 4) cli::write_bc_file(argv[3], bc);
  cli::maybe_dump_bc(bc);
  cli::write_metrics(mt, ctx);
  bcx_ctx_free(ctx);
  bcx_index_free(ix);
  bcx_graph_free(g);
  return 0;
}
