// Shared helpers of the QBC / MHLQBC / RQ executables.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "bcx.h"

namespace cli {

[[noreturn]] inline void die(const std::string &what) {
  std::cerr << "error: " << what << ": " << bcx_last_error() << std::endl;
  std::exit(1);
}

inline void check(int rc, const char *what) {
  if (rc < 0) die(what);
}

inline int env_int(const char *name, int dflt) {
  const char *v = std::getenv(name);
  return (v && *v) ? std::atoi(v) : dflt;
}

inline int parse_device(const char *s) {
  try {
    return std::stoi(s);
  } catch (...) {
    std::cerr << "error: device id must be an integer: " << s << std::endl;
    std::exit(1);
  }
}

// Knobs that the reference fixes at compile time come from the environment so
// the positional command line stays identical: BCX_MODE=plain|batched|sweep
// (default sweep), BCX_TCHUNK, BCX_BP_THREADS, BCX_BATCH, BCX_SWEEP_BATCH;
// BCX_PROFILE=1 launches kernel by kernel with events between phases, so
// that the metrics line carries kernel times.
inline int apply_env_tuning(bcx_ctx *ctx) {
  if (env_int("BCX_PROFILE", 0)) check(bcx_ctx_set_profiling(ctx, 1), "bcx_ctx_set_profiling");
  bcx_tuning tn;
  check(bcx_ctx_get_tuning(ctx, &tn), "bcx_ctx_get_tuning");
  tn.tchunk = env_int("BCX_TCHUNK", tn.tchunk);
  tn.bp_threads = env_int("BCX_BP_THREADS", tn.bp_threads);
  tn.batch = env_int("BCX_BATCH", tn.batch);
  tn.sweep_batch = env_int("BCX_SWEEP_BATCH", tn.sweep_batch);
  check(bcx_ctx_set_tuning(ctx, &tn), "bcx_ctx_set_tuning");
  const std::string m = std::getenv("BCX_MODE") ? std::getenv("BCX_MODE") : "sweep";
  if (m == "plain") return BCX_MODE_PLAIN;
  if (m == "batched") return BCX_MODE_BATCHED;
  return BCX_MODE_SWEEP;
}

// Index build flags: the label rows are left to the device (computed by the
// context) unless an image is to be saved, which needs them on the host.
inline uint32_t index_flags() {
  const char *save = std::getenv("BCX_INDEX_SAVE");
  return (save && *save) ? 0u : (uint32_t)BCX_INDEX_DEVICE_LABELS;
}

// The MHL index: loaded from an image when BCX_INDEX_LOAD names one (a
// restart without rebuilding; the context refuses an image of another
// graph), else built; saved when BCX_INDEX_SAVE names a path.
inline bcx_index *mhl_index(const bcx_graph *g) {
  bcx_index *ix = nullptr;
  const char *load = std::getenv("BCX_INDEX_LOAD");
  if (load && *load)
    check(bcx_index_load(load, &ix), "loading index");
  else
    check(bcx_index_build_flags(g, nullptr, nullptr, index_flags(), &ix), "building index");
  const char *save = std::getenv("BCX_INDEX_SAVE");
  if (save && *save) check(bcx_index_save(ix, save), "saving index");
  return ix;
}

// "vertexID  BCValue" then "id value" with 2 decimals (QBC.cu:614-626).
inline void write_bc_file(const std::string &path, const std::vector<double> &bc) {
  std::ofstream ofs;
  ofs.precision(2);
  ofs.setf(std::ios::fixed, std::ios::floatfield);
  ofs.open(path, std::ios::out);
  if (!ofs) {
    std::cerr << "error: cannot write " << path << std::endl;
    std::exit(1);
  }
  ofs << "vertexID  BCValue" << std::endl;
  for (size_t i = 0; i < bc.size(); ++i) ofs << i + 1 << " " << bc[i] << std::endl;
  ofs.close();
  std::cout << "The BC value has been save: " << path << std::endl;
}

// Optional full-precision dump for parity checks (BCX_BC_DUMP=path): raw
// little-endian doubles, one per vertex in file-id order.
inline void maybe_dump_bc(const std::vector<double> &bc) {
  const char *p = std::getenv("BCX_BC_DUMP");
  if (!p || !*p) return;
  FILE *f = std::fopen(p, "wb");
  if (!f || std::fwrite(bc.data(), sizeof(double), bc.size(), f) != bc.size()) {
    std::cerr << "error: cannot write " << p << std::endl;
    std::exit(1);
  }
  std::fclose(f);
}

// Source range of a partial run (BCX_SOURCES=begin:end, 0-based vertex ids,
// half-open): the BC file and dump then hold that range's share of BC, and
// the shares of ranges that cover every vertex once sum to the full BC, so a
// long run can be split into ranges and resumed from the last one finished.
// Returns false (all sources) when unset.
inline bool source_range(int32_t n, std::vector<int32_t> &src) {
  const char *v = std::getenv("BCX_SOURCES");
  if (!v || !*v) return false;
  long long a = -1, b = -1;
  char tail = 0;
  if (std::sscanf(v, "%lld:%lld%c", &a, &b, &tail) != 2 || a < 0 || b < a || b > n) {
    std::cerr << "error: BCX_SOURCES must be begin:end with 0 <= begin <= end <= " << n << ", got " << v
              << std::endl;
    std::exit(1);
  }
  src.resize((size_t)(b - a));
  for (long long i = a; i < b; ++i) src[(size_t)(i - a)] = (int32_t)i;
  return true;
}

// BC of every source, or of a BCX_SOURCES range (`range`, parsed before the
// index is built so that a bad value fails early); returns the sources used.
inline int64_t compute_bc(bcx_ctx *ctx, int mode, int32_t n, const std::vector<int32_t> *range,
                          std::vector<double> &bc, uint32_t &ovf) {
  bc.assign((size_t)n, 0.0);
  ovf = 0;
  if (range) {
    check(bcx_bc_sources(ctx, mode, range->data(), (int32_t)range->size(), bc.data(), &ovf), "computing BC");
    return (int64_t)range->size();
  }
  check(bcx_full_bc(ctx, mode, bc.data(), &ovf), "computing BC");
  return n;
}

inline const char *mode_name(int mode) {
  return mode == BCX_MODE_PLAIN ? "plain" : mode == BCX_MODE_BATCHED ? "batched" : "sweep";
}

// Machine-readable metrics beside the reference's stdout lines: with
// BCX_METRICS=path, one JSON object per run is appended to that file (phase
// seconds, sources or queries per second, overflow, kernel times).
struct Metrics {
  std::string tool, graph;
  int32_t n = 0;
  int64_t m = 0;
  int device = 0, mode = BCX_MODE_SWEEP;
  int64_t work = 0;          // sources (BC) or query vertices (RQ)
  const char *unit = "sources";
  double index_s = 0, context_s = 0, compute_s = 0, total_s = 0;
  uint32_t overflow = 0;
};
inline void write_metrics(const Metrics &mt, const bcx_ctx *ctx) {
  const char *p = std::getenv("BCX_METRICS");
  if (!p || !*p) return;
  bcx_timing tm{};
  if (ctx) bcx_ctx_timing(ctx, &tm);
  std::string graph;  // the path as a JSON string
  for (char c : mt.graph) {
    if ((unsigned char)c < 0x20) {
      char u[8];
      std::snprintf(u, sizeof u, "\\u%04x", (unsigned)(unsigned char)c);
      graph += u;
      continue;
    }
    if (c == '"' || c == '\\') graph += '\\';
    graph += c;
  }
  FILE *f = std::fopen(p, "a");
  if (!f) {
    std::cerr << "error: cannot append to " << p << std::endl;
    std::exit(1);
  }
  std::fprintf(f,
               "{\"tool\": \"%s\", \"graph\": \"%s\", \"n\": %d, \"m\": %lld, \"device\": %d, "
               "\"mode\": \"%s\", \"%s\": %lld, \"index_s\": %.6f, \"context_s\": %.6f, "
               "\"compute_s\": %.6f, \"total_s\": %.6f, \"%s_per_s\": %.3f, \"overflow\": %s, "
               "\"kernel_ms\": {\"scan\": %.3f, \"dep\": %.3f, \"reduce\": %.3f}}\n",
               mt.tool.c_str(), graph.c_str(), mt.n, (long long)mt.m, mt.device, mode_name(mt.mode), mt.unit,
               (long long)mt.work, mt.index_s, mt.context_s, mt.compute_s, mt.total_s, mt.unit,
               mt.compute_s > 0 ? (double)mt.work / mt.compute_s : 0.0, mt.overflow ? "true" : "false",
               tm.scan_ms, tm.dep_ms, tm.reduce_ms);
  std::fclose(f);
}

inline double seconds_since(std::chrono::high_resolution_clock::time_point t0) {
  return std::chrono::duration_cast<std::chrono::duration<double>>(
             std::chrono::high_resolution_clock::now() - t0)
      .count();
}

}  // namespace cli
