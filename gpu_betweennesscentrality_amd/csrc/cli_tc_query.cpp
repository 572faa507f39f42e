// TC_query -i index_file -q query_file -o output_file — the reference's
// TreeCount query tool (src/TreeCount/TreeCountQuery.cpp:17-60, QueryTest at
// src/TreeCount/TreeDecomp.cpp:1053-1088): reads "num_queries" then that many
// 1-based vertex pairs, and writes "num_queries" then the number of shortest
// paths of each pair, one per line.  Counts are the full uint64 values (the
// reference's uint32 counts wrap above 2^32 - 1).
#include <getopt.h>

#include "cli_common.hpp"

int main(int argc, char *argv[]) {
  std::string index, queries, answers;
  int opt;
  while ((opt = getopt(argc, argv, "i:q:o:h")) != -1) {
    if (opt == 'i') index = optarg;
    else if (opt == 'q') queries = optarg;
    else if (opt == 'o') answers = optarg;
    else {
      std::printf("TC_query -i index_file -q query_file -o output_file\n");
      return opt == 'h' ? 0 : 1;
    }
  }
  if (index.empty() || queries.empty() || answers.empty()) {
    std::printf("Error: Missing filename\n");
    return 1;
  }
  bcx_index *ix = nullptr;
  cli::check(bcx_index_load(index.c_str(), &ix), "reading index");
  bcx_index_stats st;
  bcx_index_stats_get(ix, &st);
  FILE *fin = std::fopen(queries.c_str(), "r");
  if (!fin) {
    std::fprintf(stderr, "cannot open %s\n", queries.c_str());
    return 1;
  }
  int num = 0;
  if (std::fscanf(fin, "%d", &num) != 1 || num < 0) {
    std::printf("wrong input: num_queries\n");
    return 1;
  }
  std::vector<std::pair<int, int>> qs((size_t)num);
  for (auto &q : qs)
    if (std::fscanf(fin, "%d %d", &q.first, &q.second) != 2 || q.first < 1 || q.second < 1 ||
        q.first > st.n || q.second > st.n) {
      std::printf("wrong input: no enough query data\n");
      return 1;
    }
  std::fclose(fin);
  std::vector<uint64_t> counts((size_t)num);
  const auto t0 = std::chrono::high_resolution_clock::now();
  for (size_t i = 0; i < qs.size(); ++i) {
    int32_t d;
    uint32_t h;
    if (bcx_index_query(ix, qs[i].first - 1, qs[i].second - 1, &d, &counts[i], &h) < 0)
      cli::check(-1, "query");
  }
  std::printf("Query costs %f micro seconds in average\n",
              num ? 1e6 * cli::seconds_since(t0) / num : 0.0);
  FILE *fout = std::fopen(answers.c_str(), "w");
  if (!fout) {
    std::fprintf(stderr, "cannot open %s\n", answers.c_str());
    return 1;
  }
  std::fprintf(fout, "%d\n", num);
  for (uint64_t c : counts) std::fprintf(fout, "%llu\n", (unsigned long long)c);
  std::fclose(fout);
  bcx_index_free(ix);
  return 0;
}
