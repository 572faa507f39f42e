// bcx_dev.hpp — device-side definitions shared by the HIP translation units
// of the bcx library: device views of the index and graph, batch and sweep
// workspaces, the context object behind the C ABI, and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "bcx.h"
#include "bcx_host.hpp"
#include "bcx_internal.hpp"

#define HIP_OK(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      throw bcx::Error{e_ == hipErrorOutOfMemory ? BCX_ENOMEM : BCX_EDEVICE,                 \
                       std::string(#x) + ": " + hipGetErrorString(e_)};                      \
  } while (0)

namespace bcx_dev {

constexpr int kInfDev = 999999999;
constexpr int kWave = 64;

struct DevIndex {
  int n;
  int levels;
  int64_t arcs;
  const int *depth;       // n, preorder
  const int64_t *row_off; // n + 1
  const int *dis;         // entries + 1
  const uint64_t *cnt;    // entries + 1
  const uint32_t *hop;    // entries + 1
  int64_t none;           // the entry past the labels: distance 0x7fffffff,
                          // count 0, hop 0 (a load target that matches nothing)
  const int *rmq;         // levels * n: rmq[k*n + i] = min depth over [i, i + 2^k)
  const int *grow;        // n + 1, CSR in preorder ids
  const int *gcol;        // arcs
  const int *gw;          // arcs
};

// Batch bookkeeping lives in device memory so that one captured graph can be
// replayed for every batch: src list, cursor (first list index of the batch),
// total list length.
struct DevBatch {
  const int *srcs;
  int *cursor;
  const int *total;
};

struct DevWork {
  int *d;          // slots * n
  uint64_t *sigma; // slots * n
  uint32_t *hop;   // slots * n
  int *order;      // slots * n
  int *lcnt;       // slots * (n + 2): layer bins when they exceed LDS
  double *delta;   // slots * n
};

// Tree-decomposition view for the bag sweep (all ids in tree preorder).
struct DevTree {
  int n;
  int height;
  int dbits;               // width of the distance field of the sweep's record
                           // words (every shortest distance is below 2^dbits)
  int64_t bags;            // total bag entries B
  int64_t bslots;          // backward staging slots (with padding)
  const int *bnode_slot;   // n: slot of a node's last backward piece, -1 without descendants
  const int *bslot_node;   // bslots: the node whose last piece a slot holds, or -1
  const int *subtree;      // n
  const int *bag_off;      // n + 1, entries ordered by ancestor depth
  const int *bag_v;        // B: ancestor
  const int *bag_len;      // B: restricted shortcut length
  const uint64_t *bag_cnt; // B: restricted shortest-path count
  const double *bag_cntf;  // B: the same count as a double
  const int *ibag_off;     // n + 1: entries whose ancestor is this node
  const int *ibag_x;       // B: the descendant owning the entry; sign bit set
                           // when that descendant has no descendants
  const int *ibag_e;       // B: the entry index
  const int *ibag_pos;     // B: entry index -> its position in the inverse bags
  const int *lvl_off;      // height + 1
  const int *lvl_v;        // n: nodes grouped by depth
  const int *blvl_off;     // height + 1: the same for nodes with descendants
  const int *blvl_v;
  // Tight bag pairs (a, b of bag(z) with l(a,z) + l(z,b) == l(a,b)), grouped
  // by the depth of z: pair q moves share frac[q] of shortcut (a,b)'s flow
  // (entry tp_ab[q] of bag(b)) onto z and onto z's entries for a and b.
  int64_t pairs;
  const int *lvl_pair_off; // height + 1
  const int *tp_ab;        // P
  const double *tp_frac;   // P: c(a,z) c(z,b) / c(a,b)
  const int *zp_off;       // n + 1 (by z in preorder) into zp_pair
  const int *zp_pair;      // P
  const int *ep_off;       // B + 1: pairs touching each bag entry
  const int *ep_pair;      // 2P
  // Staging slots: a node's bag / inverse-bag entries cut into pieces of at
  // most kStageEntries, one slot per piece holding the node's fields and the
  // piece's entries in fixed places, so a wavefront stages 64 pieces with one
  // round of coalesced loads.  Level k's nodes are split over the 16
  // wavefronts of a workgroup by balanced work, a node's pieces consecutive;
  // wavefront w's i-th slot is st_off[k] + w * chunk + i, chunk =
  // (st_off[k+1] - st_off[k]) / 16, and the wavefront has st_cnt[16 k + w]
  // slots.  Unused entry places name node 0 with length kSweepNoLen (never
  // minimal or tight) and count 0.
  // A level whose pieces are split over the wavefronts in contiguous runs
  // (lvl_nseg[k] > 0) leaves per-node partial results (segments) in LDS,
  // wavefront w's first one numbered st_seg[16 k + w], merged after the level.
  // In a level split by node, the last st_nar[16 k + w] (backward:
  // bst_nar) slots of wavefront w are narrow: nodes of one piece with at most
  // kNarrowEntries entries, whose steps load only that many rows; the
  // st_n3 (bst_n3) slots before them hold nodes of exactly three entries.
  const int *st_off;          // height + 1 (forward: all nodes)
  const int *st_cnt;          // 16 * height
  const int *st_seg;          // 16 * height
  const int *st_nar;          // 16 * height
  const int *st_n3;           // 16 * height
  const int *lvl_nseg;        // height: segments of a level split in runs, or 0
  const int *bst_off;         // height + 1 (backward: nodes with descendants)
  const int *bst_cnt;         // 16 * height
  const int *bst_seg;         // 16 * height
  const int *bst_nar;         // 16 * height
  const int *bst_n3;          // 16 * height
  const int *blvl_nseg;       // height
  const int4 *fs_node;        // {v, entries in the piece, 0, subtree[v] | count-1
                              // mask << 24 | last piece << 29 | first << 30}:
                              // mask bit r set when entry r's count is 1
  const int4 *fs_u;           // ancestors of the piece's entries 0..3
  const int4 *fs_len;         // their shortcut lengths
  const ulonglong2 *fs_cnt;   // 2 per slot: counts of entries 0..1, 2..3
  const int4 *bs_node;        // {v, entries in the piece, 0, leaf mask | first
                              // piece << 4 | last piece << 5 | count-1 mask
                              // << 8}: leaf-mask bit r set when entry r's
                              // descendant has none
  const int4 *bs_x;           // descendants of the piece's entries 0..3
  const int4 *bs_e;           // their entry indices
  const int4 *bs_len;         // their shortcut lengths
  const double2 *bs_cntf;     // 2 per slot: counts of entries 0..1, 2..3
};

constexpr int kStageEntries = 4;
// Default bag-sweep launch: tiles of 64 sources per compute unit (a full BC
// of a few hundred thousand sources is one launch), and the RQ engine's
// steps (each admits requests at its start).
constexpr int kSweepTilesPerCu = 32;
constexpr int kRqTilesPerCu = 8;
// Accelerator dies (XCDs) of an unpartitioned MI355X: the state-slot pool is
// split among them (a device of one die keeps one pool; other die counts do
// not pool).  Overflow-flag bit set by a workgroup that found no free slot.
constexpr int kSweepDies = 8;
constexpr int kSweepNoSlot = 2;
constexpr int kNarrowEntries = 2;
// Length of unused staged places: a distance plus it exceeds every true
// distance (below 999999999) without overflowing, so such an entry is never
// minimal or tight.
constexpr int kSweepNoLen = (1 << 30) - 1;
// Per-node partial results a level split in runs may leave (LDS space).
constexpr int kSweepMaxSegments = 48;

// Per-tile state of the sweep: one record per (tile, node) with the lanes'
// values, per-tile sums [tile][node] and [tile][entry].
struct SweepWork {
  char *state;     // per (state slot, node) record: distance[64], then 64
                   // 8-byte slots (path count, then dependency quotient); a
                   // launch's workgroup w runs its tiles in slot w
  double *node_t;  // [tile][bslots]: node flow summed over the tile's lanes, at
                   // the slot of the node's last backward piece (tree.bnode_slot)
  double *phi_t;   // [tile][B]: shortcut flow summed over the tile's lanes,
                   // in inverse-bag order (ibag_e gives the entry)
  double *node;    // [n]: node flow summed over tiles
  double *phi;     // [B]: shortcut flow summed over tiles
  double *interior;// [n]: flow through shortcut interiors (after unpacking)
  double *pf;      // [P]: flow share of each tight pair
  int *tile_src;   // [job tiles * 64] sources (tree ids) of every tile of a job
  int *cand;       // candidate lists of the ascending pass, every tile of a job
  int *cand_off;   // [job tiles + 1]
  int64_t cand_cap;
  int64_t src_cap; // tiles tile_src and cand_off hold
  int *slot_busy;  // [slot]: 1 while a workgroup holds the slot (pooled launches)
  int *tile_slot;  // [tile]: state slot that ran the launch's tile
  int slots;       // state slots allocated
  int pool;        // per launch: slots per accelerator die when the launch has
                   // more tiles than slots and draws on them as a pool, else 0
  int dies;        // accelerator dies the pool is split among (1 or kSweepDies)
  int tile0;       // job tile of a launch's tile 0 (per launch)
};

// Saturating count arithmetic: ~0 marks "overflowed" and is absorbing under
// products with nonzero counts and under sums, so a result overflowed iff its
// final count is ~0 (candidates that lose the minimum do not matter).
__device__ __forceinline__ uint64_t mul_sat(uint64_t a, uint64_t b) {
  if (((a | b) >> 32) == 0) return a * b;
  if (a == 0 || b == 0) return 0;
  if (a == ~0ull || b == ~0ull || __umul64hi(a, b) != 0) return ~0ull;
  return a * b;
}

__device__ __forceinline__ uint64_t add_sat(uint64_t a, uint64_t b) {
  const uint64_t r = a + b;
  return r < a ? ~0ull : r;
}

__device__ __forceinline__ int wave_min(int x) {
  for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, kWave));
  return x;
}
__device__ __forceinline__ int wave_max(int x) {
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, kWave));
  return x;
}

__device__ __forceinline__ double wave_sum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

template <class T>
T *dev_upload(const std::vector<T> &h) {
  T *p = nullptr;
  const size_t bytes = std::max<size_t>(1, h.size() * sizeof(T));
  HIP_OK(hipMalloc(&p, bytes));
  if (!h.empty()) HIP_OK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

}  // namespace bcx_dev

struct bcx_ctx {
  int device = 0;
  int cus = 0;  // compute units of the device
  hipStream_t stream = nullptr;
  bcx_dev::DevIndex ix{};
  std::vector<void *> owned;          // index + graph buffers
  std::vector<int32_t> perm, iperm;   // preorder <-> vertex
  bcx_tuning tn{64, 32, 256, 4096, 16384, BCX_SWEEP_FORM_AUTO, 1024, 0};
  // batch workspace
  int ws_slots = 0;
  bcx_dev::DevWork ws{};
  int *d_srcs = nullptr;  // label-scan source list (srcs_cap entries)
  int64_t srcs_cap = 0;
  int *d_cursor = nullptr;
  int *d_total = nullptr;
  int *d_ovf = nullptr;
  double *d_bc = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_batch = 0;
  bcx_tuning graph_tn{};
  bool profiling = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bcx_timing timing{};
  // bag-sweep form: tree arrays (device) and host copies used for planning
  bcx_dev::DevTree tree{};
  std::vector<int32_t> h_parent, h_depth, h_subtree, h_lvl_off, h_lvl_pair_off;
  int sweep_tiles = 0;  // tiles per launch the flow sums are allocated for
  int sweep_alloc = 0;  // state slots allocated
  int sweep_slots = 0;  // state slots the current launch shape uses (<= sweep_alloc)
  int sweep_pool = 0;   // slots per die when launches may hold more tiles than slots
  int occ_one = 1, occ_two = 2;  // resident workgroups per compute unit of each tile-kernel form
  int dies = 1;                  // accelerator dies (XCCs) of the device
  bcx_dev::SweepWork sw{};
  // A sweep job's plan on the host (tile sources, ascending-pass candidate
  // lists); it stays alive until the next job starts, which first waits for
  // the stream, so the uploads from it may run asynchronously.
  std::vector<int32_t> plan_src, plan_cand, plan_cand_off;
  std::vector<int32_t> plan_pre;  // the sorted sources of the plan in the job arrays
  // The launch sequence of the last sweep job, captured once and replayed
  // while the job's shape and buffers stay the same.
  hipGraphExec_t sweep_exec = nullptr;
  std::vector<int64_t> sweep_exec_key;
  // QBC H2H distance labels and separators, when the index had them at
  // creation.  The labels are interleaved by groups of 64 consecutive
  // preorder ids: entry k of node t at d_goff[t / 64] + 64 k + t % 64 (a
  // group's rows padded to its longest), so that a wavefront of consecutive
  // targets reads one entry position as one contiguous 256-byte row.
  const int *d_gdis = nullptr;
  const int64_t *d_goff = nullptr;
  const int64_t *d_sep_off = nullptr;
  const int *d_sep = nullptr;
  const int *d_iperm = nullptr;  // vertex id -> preorder id
  // P2H branch tables (see bcx::Index): per-node slot, per-slot base, offsets, lists.
  const int *d_p2h_slot = nullptr;
  const int64_t *d_p2h_base = nullptr;
  const int64_t *d_p2h_off = nullptr;
  const int *d_p2h = nullptr;

  ~bcx_ctx() {
    if (hipSetDevice(device) != hipSuccess) return;
    if (stream) (void)hipStreamSynchronize(stream);
    free_ws();  // destroys graph_exec once and clears it
    free_sweep();  // destroys sweep_exec
    if (d_srcs) (void)hipFree(d_srcs);
    for (void *p : owned) (void)hipFree(p);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void free_sweep();
  void ensure_sweep(int slots, int tiles);
  void free_ws() {
    if (graph_exec) {
      (void)hipGraphExecDestroy(graph_exec);
      graph_exec = nullptr;
    }
    void *bufs[] = {ws.d, ws.sigma, ws.hop, ws.order, ws.lcnt, ws.delta};
    for (void *p : bufs)
      if (p) (void)hipFree(p);
    ws = bcx_dev::DevWork{};
    ws_slots = 0;
  }
  // Room for `count` label-scan sources; the captured batch graph names the
  // list, so it is dropped when the list moves.
  void ensure_srcs(int64_t count) {
    if (count <= srcs_cap && d_srcs) return;
    if (stream) HIP_OK(hipStreamSynchronize(stream));
    if (graph_exec) {
      (void)hipGraphExecDestroy(graph_exec);
      graph_exec = nullptr;
    }
    if (d_srcs) HIP_OK(hipFree(d_srcs));
    d_srcs = nullptr;
    srcs_cap = 0;
    const int64_t cap = std::max<int64_t>({count, (int64_t)ix.n, 1});
    HIP_OK(hipMalloc(&d_srcs, (size_t)cap * 4));
    srcs_cap = cap;
  }
  static size_t slot_bytes(int n) {
    return (size_t)n * (4 + 8 + 4 + 4 + 8) + ((size_t)n + 2) * 4;
  }
  // The label-scan workspace for `slots` sources.  When device memory runs
  // out and the bag sweep's buffers are held, they are released (they are
  // re-allocated on the sweep's next use) and the allocation is tried again,
  // as the sweep releases this workspace in the other direction.
  void ensure_ws(int slots) {
    if (slots <= ws_slots) return;
    try {
      alloc_ws(slots);
    } catch (const bcx::Error &e) {
      if (e.code != BCX_ENOMEM || sweep_alloc == 0) throw;
      (void)hipGetLastError();  // clear the failed allocation's error
      if (stream) HIP_OK(hipStreamSynchronize(stream));
      free_sweep();
      alloc_ws(slots);
    }
  }
  void alloc_ws(int slots) {
    free_ws();
    const size_t n = (size_t)ix.n;
    HIP_OK(hipMalloc(&ws.d, n * slots * 4));
    HIP_OK(hipMalloc(&ws.sigma, n * slots * 8));
    HIP_OK(hipMalloc(&ws.hop, n * slots * 4));
    HIP_OK(hipMalloc(&ws.order, n * slots * 4));
    HIP_OK(hipMalloc(&ws.lcnt, (n + 2) * slots * 4));
    HIP_OK(hipMalloc(&ws.delta, n * slots * 8));
    ws_slots = slots;
  }
};

namespace bcx_dev {
// Sorts tree ids in [0, n) ascending, repeats kept: a counting sort when the
// list is long against n (a full BC's 300k sources: 4 ms instead of 22 ms of
// comparison sorting), else std::sort.
inline void sort_tree_ids(std::vector<int32_t> &ids, int32_t n) {
  if ((int64_t)ids.size() * 8 < (int64_t)n) {
    std::sort(ids.begin(), ids.end());
    return;
  }
  std::vector<int32_t> start((size_t)n + 1, 0);
  for (int32_t x : ids) ++start[(size_t)x + 1];
  for (int32_t i = 0; i < n; ++i) start[(size_t)i + 1] += start[(size_t)i];
  std::vector<int32_t> out(ids.size());
  for (int32_t x : ids) out[(size_t)start[(size_t)x]++] = x;
  ids.swap(out);
}
// Bag-sweep full BC over `count` sources (vertex ids, repeats allowed),
// added into bc_inout.
void run_sweep(bcx_ctx *c, const int32_t *sources, int32_t count, double *bc_inout,
               uint32_t *overflow);
// Distance and path-count rows of the bag sweep's forward pass for `count`
// sources (vertex ids): d[i * n + v], sigma[i * n + v] for sources[i].
// Returns 1 if some count saturated.
int sweep_query(bcx_ctx *c, const int32_t *sources, int32_t count, int32_t *d, uint64_t *sigma);
// Pieces of the bag sweep, for callers that keep flows per batch.
// Tiles per launch for `count` sources (tuning batch, capped by device
// memory); allocates the sweep buffers: state slots for the launch's
// resident workgroups and flow sums for that many tiles.
int sweep_tiles_for(bcx_ctx *c, int32_t count);
// Tile-kernel form of a launch of nt tiles with the allocated slots.
int sweep_launch_form(const bcx_ctx *c, int nt);
// Resident workgroups per compute unit of each tile-kernel form (ctx occ_*).
void sweep_occupancy(bcx_ctx *c);
// One launch over the sorted tree ids pre[0 .. nsrc) (at most sweep_tiles
// tiles of 64): node flows into node_out[n], shortcut flows into phi_out[B],
// added when `accumulate`, else stored.  Sets ctx d_ovf on count overflow.
void sweep_launch(bcx_ctx *c, const int32_t *pre, int32_t nsrc, double *node_out,
                  double *phi_out, bool accumulate);
// ctx sw.node / sw.phi = sum of nb per-batch slices ([nb][n], [nb][B]) in order.
void sweep_sum_batches(bcx_ctx *c, const double *nodeb, const double *phi_b, int nb);
// Unpack sw.node / sw.phi and add 0.5 * (node flow + interior flow) to bc_dev[n]
// (tree ids).
void sweep_finish(bcx_ctx *c, double *bc_dev);
void sweep_zero(bcx_ctx *c, double *p, int64_t cnt);
// Throws if a pooled launch since the flags were last cleared found no free
// state slot (its tiles then went unswept); call after the stream is idle.
void sweep_check_slots(bcx_ctx *c);
}  // namespace bcx_dev
