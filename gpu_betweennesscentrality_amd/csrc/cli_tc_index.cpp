// TC_index -g graph_file -o output_file — the reference's TreeCount index
// tool (src/TreeCount/TreeCount.cpp:13-62): builds the count-label index of a
// graph and writes it to a file (this library's binary index image, read
// back by TC_query and by bcx_index_load).
#include <getopt.h>

#include "cli_common.hpp"

int main(int argc, char *argv[]) {
  std::string graph, out;
  int opt;
  while ((opt = getopt(argc, argv, "g:o:h")) != -1) {
    if (opt == 'g') graph = optarg;
    else if (opt == 'o') out = optarg;
    else {
      std::printf("TC_index -g graph_file -o output_file\n");
      return opt == 'h' ? 0 : 1;
    }
  }
  if (graph.empty() || out.empty()) {
    std::printf("Error: Missing filename\n");
    return 1;
  }
  const auto t0 = std::chrono::high_resolution_clock::now();
  bcx_graph *g = nullptr;
  cli::check(bcx_graph_read(graph.c_str(), &g), "reading graph");
  bcx_index *ix = nullptr;
  cli::check(bcx_index_build_mhl(g, &ix), "building index");
  std::printf("MakeIndex is OK, time from begin: %f ms\n", 1e3 * cli::seconds_since(t0));
  cli::check(bcx_index_save(ix, out.c_str()), "writing index");
  std::printf("The index has been written to: %s\n", out.c_str());
  bcx_index_free(ix);
  bcx_graph_free(g);
  return 0;
}
