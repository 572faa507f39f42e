// GraphConstract <graph> <orderOut> [notes] — writes the min-degree order file
// that QBC reads, as Graph::GraphConstract does (src/buildIndex.cpp:759-925):
// "n", then "nodeID order" for every 0-based node id, then the notes line;
// stdout carries the reference's progress and summary lines.
#include "cli_common.hpp"

int main(int argc, char *argv[]) {
  if (argc < 3) {
    std::cerr << "usage: GraphConstract <graph> <orderOut> [notes]" << std::endl;
    return 1;
  }
  bcx_graph *g = nullptr;
  cli::check(bcx_graph_read(argv[1], &g), "reading graph");
  int32_t n;
  int64_t m;
  bcx_graph_info(g, &n, &m);
  std::vector<int32_t> order((size_t)n), widths((size_t)n / 100000 + 1);
  int64_t super_arcs = 0;
  cli::check(bcx_min_degree_order(g, order.data(), &super_arcs, widths.data()), "ordering");
  for (size_t i = 0; i * 100000 < (size_t)n; ++i)
    std::cout << "count " << i * 100000 << " , treewidth " << widths[i] << std::endl;
  std::ofstream out(argv[2]);
  if (!out) {
    std::cerr << "cannot open " << argv[2] << std::endl;
    return 1;
  }
  out << n << "\n";
  for (int32_t v = 0; v < n; ++v) out << v << " " << order[(size_t)v] << "\n";
  out << (argc >= 4 ? argv[3] : "") << std::endl;
  std::cout << "edge number " << m << " super edge number " << super_arcs << " ratio "
            << (double)super_arcs / (double)m << std::endl;
  bcx_graph_free(g);
  return out ? 0 : 1;
}
