// bcx_sweep.hip — bag-sweep form of the full-BC path (BCX_MODE_SWEEP).
//
// Same index, same results as the label-scan form, different evaluation
// order.  For a source s, every shortest s-t path has a unique up-down form
// over the decomposition: it climbs through ancestors of s along bag
// shortcuts to its highest vertex h (a prefix of s's label, TreeDecomp.cpp:
// 402-415) and descends from h to t, each step again a bag shortcut whose
// interior lies below both ends.  Hence, in tree order:
//   forward   d(s,v) = min( up-label of s at v if v is an ancestor of s,
//                           min over bag entries (u, w) of v of d(s,u) + w ),
//             sigma(s,v) = sum of the tight candidates' counts (label count,
//                           sigma(s,u) * shortcut count);
//   backward  Brandes on that shortcut DAG, deepest node first:
//             dD(v) = sum over tight entries (x has v in its bag)
//                     c * sigma(v) * (1 + dD(x)) / sigma(x)
//             (the reference's sigma_u * (1 + delta_v) / sigma_v,
//             MHLQBC.cu:457, with the shortcut multiplicity c), and the same
//             recurrence on the ascending states of s's ancestors;
//   unpack    the flow on each shortcut (summed over all sources) is split,
//             top-down, over the vertices eliminated inside it in proportion
//             to their share of the shortcut's restricted path count.
// BC(v) = 0.5 * (flow through v as a shortcut end + flow through v inside
// shortcuts), the reference's 0.5 * sum_s delta_s(v) (MHLQBC.cu:470).
// Work per (source, node) is the bag size (~3.4 on road graphs) instead of a
// label prefix (~250) plus a per-source layer sort, and every step is a sweep
// in tree order shared by the 64 sources of a tile: lanes are sources, so
// node state is [node][lane] and every access is one coalesced row.
#include <algorithm>
#include <cstring>
#include <vector>

#include "bcx_dev.hpp"

using namespace bcx_dev;

namespace {

// Lane state is one record per (tile, node): 64 distances, then 64 8-byte
// slots (768 bytes), so the rows a node reads of an ancestor or descendant
// are one contiguous region of memory.  The slot holds the path count sigma
// through the forward sweep; a node's backward step is the last reader of
// its own count and replaces it with the dependency quotient
// Q = (1 + dD) / sigma, which its ancestors then read.
constexpr int64_t kSlotOff = kWave * 4;
constexpr int64_t kRec = kWave * (4 + 8);

// Offset beyond every buffer resource below: a raw buffer store given it in a
// lane writes nothing there (out-of-range buffer stores are dropped), while
// the instruction still issues and is counted like any other.  The pipelined
// steps issue every store of a step (see wait_vm_left); lanes and steps that
// must not write use it, which costs no memory traffic.
constexpr uint32_t kDrop = 0xF0000000u;
// Word 3 of a gfx9-family buffer resource for untyped accesses (32-bit data
// format; no swizzle, no index stride).
constexpr int kBufferWord3 = 0x00020000;

struct Lane {
  int t, lane, s;
  bool valid;
  int64_t rs;
  char *tb;      // the tile's records (uniform); a record is addressed by its
                 // 32-bit offset (node * kRec), the tile spanning under kDrop bytes
  __amdgpu_buffer_rsrc_t rec;   // the tile's records as a buffer resource
  __amdgpu_buffer_rsrc_t phi;   // the tile's shortcut flows (sw.phi_t row)
  __amdgpu_buffer_rsrc_t nodef; // the tile's node flows (sw.node_t row)
  uint32_t dbits;  // distance field width of the record words (uniform)
  uint32_t dmask;  // 2^dbits - 1
  uint32_t cmax;   // largest count kept in a word, 2^(32 - dbits) - 1
};

// Tile t of a launch is tile sw.tile0 + t of the job: its sources and
// candidate lists are the job's, its flow sums the launch's.  Its records
// are those of the state slot of the workgroup that runs it.
__device__ __forceinline__ Lane lane_setup(const DevIndex &ix, const DevTree &tr, const SweepWork &sw,
                                           int t, int slot, int nsrc, int n) {
  Lane L;
  L.t = t;
  L.lane = threadIdx.x & (kWave - 1);
  const int64_t src = (int64_t)(sw.tile0 + t) * kWave + L.lane;
  L.valid = src < nsrc;
  L.s = sw.tile_src[src];  // padded with a valid source beyond nsrc
  L.rs = ix.row_off[L.s];
  L.tb = sw.state + (int64_t)slot * (n + 1) * kRec;
  L.rec = __builtin_amdgcn_make_buffer_rsrc(L.tb, 0, (int)((uint32_t)(n + 1) * (uint32_t)kRec), kBufferWord3);
  L.phi = __builtin_amdgcn_make_buffer_rsrc(sw.phi_t + (int64_t)t * tr.bags, 0, (int)(uint32_t)(tr.bags * 8),
                                            kBufferWord3);
  L.nodef = __builtin_amdgcn_make_buffer_rsrc(sw.node_t + (int64_t)t * tr.bslots, 0,
                                              (int)(uint32_t)(tr.bslots * 8), kBufferWord3);
  L.dbits = (uint32_t)tr.dbits;
  L.dmask = (1u << L.dbits) - 1u;
  L.cmax = 0xffffffffu >> L.dbits;
  return L;
}

// Read-only data at a wave-uniform index, through the constant address space
// so that it becomes a scalar load (scalar cache, no vector instructions).
template <class T>
__device__ __forceinline__ T cld(const T *p, int64_t i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}

// Raw buffer stores (32- and 64-bit) at a byte offset of a resource.
__device__ __forceinline__ void st32(__amdgpu_buffer_rsrc_t r, uint32_t off, int x) {
  __builtin_amdgcn_raw_buffer_store_b32(x, r, (int)off, 0, 0);
}
__device__ __forceinline__ void st64(__amdgpu_buffer_rsrc_t r, uint32_t off, uint64_t x) {
  typedef int v2i __attribute__((ext_vector_type(2)));
  const v2i v = {(int)(uint32_t)x, (int)(uint32_t)(x >> 32)};
  __builtin_amdgcn_raw_buffer_store_b64(v, r, (int)off, 0, 0);
}

// Record offset of node v (uniform).  A row access is the uniform tile base
// plus one 32-bit per-lane offset, so each costs one full-rate add.
__device__ __forceinline__ uint32_t roff(int v) { return (uint32_t)v * (uint32_t)kRec; }
__device__ __forceinline__ int &SD(const Lane &L, uint32_t r) {
  return *(int *)(L.tb + (r + (uint32_t)L.lane * 4u));
}
__device__ __forceinline__ uint64_t &SS(const Lane &L, uint32_t r) {
  return *(uint64_t *)(L.tb + (r + (uint32_t)kSlotOff + (uint32_t)L.lane * 8u));
}
__device__ __forceinline__ void setQ(const Lane &L, uint32_t r, double q) {
  SS(L, r) = __builtin_bit_cast(uint64_t, q);
}

// A distance word holds the lane's distance in its low dbits bits and the
// lane's path count in the bits above, when the count is between 1 and
// 2^(32 - dbits) - 1; a count field of 0 says the count is in the record's
// 8-byte slot.  dbits covers the largest possible distance (twice the
// farthest root distance, below 2^30), so on road graphs, with 17 to 20 bits
// of distance, counts up to several thousand travel with the distance: the
// count slot is neither written nor read for them, also where ties are
// common (a third to two thirds of the lanes on weights of about 100).
// Length given to entries past the end of a bag: a candidate distance plus it
// exceeds every true distance (below 999999999) without overflowing, so such
// an entry is never minimal or tight.  Selecting it on the scalar length keeps
// the vector arithmetic on loaded rows free of branches (a uniform select of a
// loaded value becomes a branch that waits for the load).
constexpr int kNoLen = kSweepNoLen;
static_assert(kNoLen >= (1 << 30) - 1 && kNoLen < (1 << 30), "a distance plus kNoLen stays below 2^31");
__device__ __forceinline__ int dist_of(const Lane &L, int w) { return w & (int)L.dmask; }
// The count field of a word: the count, or 0 where the count is in the slot.
__device__ __forceinline__ uint32_t cnt_of(const Lane &L, int w) { return (uint32_t)w >> L.dbits; }
__device__ __forceinline__ bool cnt_in_word(const Lane &L, uint64_t c) { return c - 1 < (uint64_t)L.cmax; }
__device__ __forceinline__ int pack_word(const Lane &L, int d, uint64_t c) {
  return d | (int)((cnt_in_word(L, c) ? (uint32_t)c : 0u) << L.dbits);
}
__device__ __forceinline__ void put_dist_count(const Lane &L, uint32_t r, int d, uint64_t c) {
  SD(L, r) = pack_word(L, d, c);
  if (!cnt_in_word(L, c)) SS(L, r) = c;
}

// Stores inside the pipelined steps are issued unconditionally, with lanes
// (or whole steps) that must not write given the kDrop offset.  Loads and
// stores share one completion counter, so the wait for the next step's rows
// (issued before these stores) can then leave exactly these stores
// outstanding; a store under a condition would make that wait drain them.
// put: the lane's node result (distance word, and the count where it is not
// 1) goes to record r.
__device__ __forceinline__ void put_dist_count_all(const Lane &L, bool put, uint32_t r, int d, uint64_t c) {
  st32(L.rec, put ? r + (uint32_t)L.lane * 4u : kDrop, pack_word(L, d, c));
  st64(L.rec, (put & !cnt_in_word(L, c)) ? r + (uint32_t)kSlotOff + (uint32_t)L.lane * 8u : kDrop, c);
}
__device__ __forceinline__ void put_q_all(const Lane &L, bool put, uint32_t r, double q) {
  st64(L.rec, put ? r + (uint32_t)kSlotOff + (uint32_t)L.lane * 8u : kDrop, __builtin_bit_cast(uint64_t, q));
}
__device__ __forceinline__ void put_flow_all(__amdgpu_buffer_rsrc_t res, bool put, int64_t idx, double x) {
  st64(res, put ? (uint32_t)idx * 8u : kDrop, __builtin_bit_cast(uint64_t, x));
}
// kCount dropped stores after a pipelined loop's first set of row loads: the
// loop body's first wait then sees as many stores after the rows on entry as
// on every later iteration, and leaves them outstanding.  The compiler
// barrier keeps the repeated stores from being merged.
template <int kCount>
__device__ __forceinline__ void sink_stores(const Lane &L) {
#pragma unroll
  for (int i = 0; i < kCount; ++i) {
    st32(L.rec, kDrop, 0);
    __asm__ volatile("" ::: "memory");
  }
}

// Waits until at most kLeft loads and stores are outstanding.  At the end of
// a pipelined step, kLeft is the next step's row loads plus this step's
// stores: everything older, including loads whose values went unused on some
// path, is then known complete, so no register of this step counts as pending
// in the next (a pending register is waited for before it is overwritten).
template <int kLeft>
__device__ __forceinline__ void wait_vm_left() {
  static_assert(kLeft < 64, "vmcnt field");
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(0x0f70 | (kLeft & 15) | ((kLeft >> 4) << 14));
  __builtin_amdgcn_sched_barrier(0);
}

// Waits for every load and store issued so far.  Used once per staged batch
// of 64 nodes: the staged fields are then known to be in registers, so the
// steps that read them need no wait of their own (which the compiler would
// otherwise place, counted conservatively, inside every step).
__device__ __forceinline__ void wait_all_vm() {
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0), expcnt and lgkmcnt not waited
  __builtin_amdgcn_sched_barrier(0);
}
// Quotient Q = (1 + dD) / sigma of a node from its word and slot.  A node
// without descendants has dD = 0 and its backward step is skipped, so its
// slot still holds sigma (read only where the word does not hold it); other
// nodes hold Q itself.  (Storing 1 / sigma from the forward pass instead
// moves the divisions there, and measured no faster.)
__device__ __forceinline__ bool slot_needed(const Lane &L, bool leaf, int w) {
  return !(leaf & (cnt_of(L, w) != 0));
}
__device__ __forceinline__ double q_from_slot(const Lane &L, bool leaf, int w, uint64_t raw) {
  const uint32_t cw = cnt_of(L, w);
  return leaf ? 1.0 / (double)(cw ? (uint64_t)cw : raw) : __builtin_bit_cast(double, raw);
}

// Slot word of record r in lanes where `need` holds, dflt in the others.  The
// load runs under the lane mask, so lanes that need nothing move no data
// through the memory pipeline, and a step where no lane needs it skips it.
// The pipelined steps wait for such loads with explicit waits (wait_vm_left),
// which hold whether or not they issued.
__device__ __forceinline__ uint64_t slot_or(bool need, uint32_t r, const Lane &L, uint64_t dflt) {
  uint64_t x = dflt;
  if (need) x = SS(L, r);
  return x;
}

// Lane moves of a double.  dpp_row: a DPP move inside rows of 16 lanes that
// writes every lane; dpp_move: one with a row mask (rows left out keep 0).
template <int kCtrl>
__device__ __forceinline__ double dpp_row(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, kCtrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), kCtrl, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_move(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, kCtrl, kRowMask, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, kCtrl, kRowMask, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over each row of 16 lanes, left in every lane of the row.
__device__ __forceinline__ double row_sum(double v) {
  v += dpp_row<0xb1>(v);   // quad_perm [1,0,3,2]
  v += dpp_row<0x4e>(v);   // quad_perm [2,3,0,1]: quad sums
  v += dpp_row<0x124>(v);  // row_ror:4
  v += dpp_row<0x128>(v);  // row_ror:8
  return v;
}

// Sum of a double over the 64 lanes of a wavefront, in one fixed order,
// returned in every lane.  All 64 lanes must be active.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v = row_sum(v);
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Cross-half and cross-row pair sums with the gfx950 lane swaps:
//   half_pair_sum(a, b): lanes 0..31 get a[i] + a[i + 32], lanes 32..63 get
//                        b[i - 32] + b[i];
//   row_pair_sum(a, b):  rows 0, 2 get a[i] + a[i + 16], rows 1, 3 get
//                        b[i - 16] + b[i].
__device__ __forceinline__ double half_pair_sum(double a, double b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  const double x = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  const double y = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
  return x + y;
}
__device__ __forceinline__ double row_pair_sum(double a, double b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi =
      __builtin_amdgcn_permlane16_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  const double x = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  const double y = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
  return x + y;
}

// Four 64-lane sums at once (a transposed reduction): every lane of row r
// (lanes 16r .. 16r + 15) receives the sum over all lanes of a_r.  Six lane
// swaps and four row steps instead of four full reductions.  All 64 lanes
// must be active.
__device__ __forceinline__ double wave_sum4(double a0, double a1, double a2, double a3) {
  const double b0 = half_pair_sum(a0, a2);  // a0 in lanes 0..31, a2 in 32..63
  const double b1 = half_pair_sum(a1, a3);  // a1 in lanes 0..31, a3 in 32..63
  return row_sum(row_pair_sum(b0, b1));     // rows 0..3: a0, a1, a2, a3
}

// Of four uniform values, the one for this lane's row.
__device__ __forceinline__ int row_pick(int lane, int x0, int x1, int x2, int x3) {
  const int r = lane >> 4;
  return r == 0 ? x0 : r == 1 ? x1 : r == 2 ? x2 : x3;
}

constexpr int kSweepThreads = 1024;
constexpr int kSweepWaves = kSweepThreads / kWave;

__device__ __forceinline__ int rl(int x, int j) { return __builtin_amdgcn_readlane(x, j); }

constexpr int kStage = kStageEntries;  // bag entries per node staged in registers

// Row loads below are issued unconditionally, from in-bounds indices (staged
// entries beyond a node's bag name node 0), and masked with a select
// afterwards: a load under a condition becomes a branch that waits for it,
// which serializes the loads of a step.

// Nodes of a level processed together by one wavefront (forward kNodesF,
// backward kNodesB) are template parameters of the tile kernel.

// This wavefront's staging slots of level k: [first, first + count).  The
// host splits each level's nodes over the 16 wavefronts by balanced work and
// keeps tree order within a wavefront, so the wavefronts of a workgroup touch
// neighbouring (ancestor-sharing) rows at about the same time.
__device__ __forceinline__ void level_block(const int *st_off, const int *st_cnt, int k, int wave,
                                            int &first, int &count) {
  const int o0 = cld(st_off, k), o1 = cld(st_off, k + 1);
  const int chunk = (o1 - o0) / kSweepWaves;
  first = o0 + wave * chunk;
  count = cld(st_cnt, k * kSweepWaves + wave);
}

// A staged piece's entry fields (kWidth of the four places of its slot).
template <int kWidth>
__device__ __forceinline__ void stage_entries(const int4 *ids, const int4 *lens, int p,
                                              int (&id)[kWidth], int (&len)[kWidth]) {
  if constexpr (kWidth == 4) {
    const int4 a = ids[p], b = lens[p];
    id[0] = a.x, id[1] = a.y, id[2] = a.z, id[3] = a.w;
    len[0] = b.x, len[1] = b.y, len[2] = b.z, len[3] = b.w;
  } else if constexpr (kWidth == 3) {
    const int4 a = ids[p], b = lens[p];
    id[0] = a.x, id[1] = a.y, id[2] = a.z;
    len[0] = b.x, len[1] = b.y, len[2] = b.z;
  } else {
    static_assert(kWidth == 2, "narrow pieces hold two entries");
    const int2 a = ((const int2 *)ids)[2 * p], b = ((const int2 *)lens)[2 * p];
    id[0] = a.x, id[1] = a.y;
    len[0] = b.x, len[1] = b.y;
  }
}

// LDS segments of a level split in runs (see the staging): per segment, two
// 64-bit words per lane and the node it belongs to.
struct Segs {
  uint64_t (*x)[kWave];
  uint64_t (*y)[kWave];
  int *node;
};

// Forward over this wavefront's block of a level (all at tree depth `depth`).
// Lane j stages piece j (a node's fields and up to kStage of its bag entries)
// with one set of loads; the wavefront then handles kNodesF pieces per
// step.  Load set 1 of a step loads the distance words of the pieces' entries;
// load set 2 the up-label (first piece, where the node is an ancestor of the
// lane's source) and the count slots of tight candidates whose distance words
// hold no count.
// The steps are pipelined: a step issues the next step's load set 1 after its
// own load set 2 and before its stores, so every wait is for loads that are
// already old (stores and loads share one completion counter).  A node's
// minimum and count are carried from piece to piece and stored by its last
// piece.  Saturating sums of saturating products do not depend on the order
// of their terms, so the count equals that of a running min-and-count.
// The wavefront's slots [first, first + mine), of at most kWidth entries each
// (kStage, or kNarrowEntries for the narrow tail of a level split by node).
template <int kNodesF, bool runs, int kWidth>  // runs: split in runs, results go to LDS segments
__device__ __forceinline__ void fwd_level_wave(const DevIndex &ix, const DevTree &tr,
                                               const Lane &L, int depth, int first, int mine,
                                               int seg, int &ovf, const Segs &sg) {
  static_assert(kWidth == kStage || (!runs && (kWidth == 3 || kWidth == kNarrowEntries)),
                "piece width");
  int cb = 0x7fffffff;  // the node's minimum and count over its earlier pieces
  uint64_t cc = 0;
  for (int base = 0; base < mine; base += kWave) {
    const int jn = min(kWave, mine - base);
    const int p = first + base + min(L.lane, jn - 1);
    const int4 nd = tr.fs_node[p];
    int sur[kWidth], slr[kWidth];
    stage_entries<kWidth>(tr.fs_u, tr.fs_len, p, sur, slr);
    wait_all_vm();
    // Shortcut counts are 1 for most entries (a flag bit); the others are
    // read as scalars when used, which keeps them out of the staged registers.
    const uint64_t *scnt = (const uint64_t *)tr.fs_cnt + 4 * (int64_t)(first + base);
    // Load set 1 of the step at j0: the staged entries' distance rows (nodes past
    // the block repeat its last node).
    // Every staged place is loaded, also past the node's bag (node 0's row,
    // a hot line): a load that may not issue would shorten every count-based
    // wait after it, and skipping them measured slower.
    auto rows1 = [&](int j0, int (&w)[kNodesF][kWidth]) {
#pragma unroll
      for (int u = 0; u < kNodesF; ++u) {
        const int j = min(j0 + u, jn - 1);
#pragma unroll
        for (int r = 0; r < kWidth; ++r) w[u][r] = SD(L, roff(rl(sur[r], j)));
      }
    };
    int w[kNodesF][kWidth];
    rows1(0, w);
    sink_stores<2 * kNodesF>(L);  // a step stores a distance and a count row per node
    for (int j0 = 0; j0 < jn; j0 += kNodesF) {
      // This step's rows have arrived once only the stores issued after them
      // (the last step's, or the sink stores) are outstanding.
      wait_vm_left<2 * kNodesF>();
      int cand[kNodesF][kWidth], best[kNodesF];
#pragma unroll
      for (int u = 0; u < kNodesF; ++u) {
        const int j = min(j0 + u, jn - 1);
        best[u] = 0x7fffffff;
#pragma unroll
        for (int r = 0; r < kWidth; ++r) {
          const int len = rl(slr[r], j);  // kNoLen in unused places
          cand[u][r] = dist_of(L, w[u][r]) + len;
          best[u] = min(best[u], cand[u][r]);
        }
      }
      // Phase order is kept as written: all candidates, then every load of
      // the step (a load under a lane mask may not issue, so the one wait
      // that follows covers all of them), then the arithmetic and stores.
      __builtin_amdgcn_sched_barrier(0);
      // Load set 2, first the up-label where v is an ancestor of the lane's
      // source.  A tile's sources are neighbours in tree order, so for nearly
      // every node no lane needs it and the wavefront skips both loads; other
      // lanes read the entry that matches nothing, so the values need no
      // select (which would wait for them here).  The bag minimum above leaves
      // the up-label out: where the up-label is smaller no candidate is tight
      // and the count rows loaded below go unused; where it ties, the tight
      // candidates are the same.
      int upd[kNodesF];
      uint64_t upc[kNodesF];
#pragma unroll
      for (int u = 0; u < kNodesF; ++u) {
        const int j = min(j0 + u, jn - 1);
        const int v = rl(nd.x, j), fl = rl(nd.w, j), size = fl & ((1 << 24) - 1);
        const bool anc = (v <= L.s) & (L.s < v + size);
        upd[u] = 0x7fffffff;
        upc[u] = 0;
        if (((fl >> 30) & 1) && __builtin_amdgcn_ballot_w64(anc)) {  // first piece only
          const int64_t o = anc ? L.rs + depth : ix.none;
          upd[u] = ix.dis[o];
          upc[u] = ix.cnt[o];
        }
      }
      // Then the count slots of tight candidates whose words hold no count
      // (m is the candidate's count in its tight lanes).
      uint64_t m[kNodesF][kWidth];
#pragma unroll
      for (int u = 0; u < kNodesF; ++u) {
        const int j = min(j0 + u, jn - 1);
#pragma unroll
        for (int r = 0; r < kWidth; ++r) {
          const uint32_t cw = cnt_of(L, w[u][r]);
          m[u][r] = slot_or((cand[u][r] == best[u]) & (cw == 0), roff(rl(sur[r], j)), L, cw);
        }
      }
      rows1(j0 + kNodesF, w);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < kNodesF; ++u) {
        const int j = min(j0 + u, jn - 1);
        const bool live = j0 + u < jn;  // false for the clamped repeats of the last piece
        const int v = rl(nd.x, j), fl = rl(nd.w, j);
        // A result starts at the node's first piece (or at the wavefront's
        // first, when split in runs) and ends at its last (or the
        // wavefront's last).
        const int pos = base + j0 + u;
        const bool pfirst = ((fl >> 30) & 1) || (runs && pos == 0);
        const bool fin = live && (((fl >> 29) & 1) || (runs && pos == mine - 1));
        int b = min(best[u], upd[u]);
        uint64_t c = upd[u] == b ? upc[u] : 0;
#pragma unroll
        for (int r = 0; r < kWidth; ++r)
          if (cand[u][r] == b)  // shortcut counts are 1 for most entries (a uniform bit)
            c = add_sat(c, ((fl >> (24 + r)) & 1) ? m[u][r] : mul_sat(m[u][r], cld(scnt, 4 * j + r)));
        // Merge with the node's earlier pieces (saturating sums do not
        // depend on how their terms are grouped).
        if (!pfirst) {
          c = cb < b ? cc : cb == b ? add_sat(c, cc) : c;
          b = min(b, cb);
        }
        if (live) {
          cb = b;
          cc = c;
        }
        if constexpr (runs) {
          if (fin) {  // uniform
            sg.x[seg][L.lane] = c;
            sg.y[seg][L.lane] = (uint32_t)b;
            if (L.lane == 0) sg.node[seg] = v;
            ++seg;
          }
          put_dist_count_all(L, false, 0, b, c);
          __asm__ volatile("" ::: "memory");  // not a dead store: the waits count it
        } else {
          put_dist_count_all(L, fin, roff(v), b, c);
          ovf |= fin & L.valid & (c == ~0ull);
        }
      }
      wait_vm_left<kNodesF * kWidth + 2 * kNodesF>();
    }
  }
}

// Backward over this wavefront's block of a level's nodes that have
// descendants (a node without descendants has dD = 0 and no flows; its Q is
// derived from its count by the readers), with the same staging by pieces
// (of the inverse bag) and kNodesB pieces per step.  Load set 1 loads the node's
// distance word and the piece's entries' distance words; load set 2 the node's
// count (lanes whose word holds none) and the tight entries' quotients; steps are
// pipelined as in the forward pass.  The node's dependency part is summed
// piece by piece; its last piece writes Q and the node flow.  The flows of a
// piece's entries and the node flows of the step are each one four-way
// transposed lane sum.
template <int kNodesB, bool runs, int kWidth>  // runs: split in runs, results go to LDS segments
__device__ __forceinline__ void bwd_level_wave(const DevTree &tr, const SweepWork &sw,
                                               const Lane &L, int first, int mine, int seg,
                                               const Segs &sg) {
  static_assert(kNodesB <= 4 && kStage == 4, "four-way lane sums");
  static_assert(kWidth == kStage || (!runs && (kWidth == 3 || kWidth == kNarrowEntries)),
                "piece width");
  double carry = 0.0;  // the node's flow sum over its earlier pieces
  for (int base = 0; base < mine; base += kWave) {
    const int jn = min(kWave, mine - base);
    const int p = first + base + min(L.lane, jn - 1);
    const int4 nd = tr.bs_node[p];
    int sxr[kWidth], slr[kWidth];
    stage_entries<kWidth>(tr.bs_x, tr.bs_len, p, sxr, slr);
    wait_all_vm();
    // Shortcut counts (1 for most entries, a flag bit) and the entries' shortcut
    // indices are read as scalars when used, not staged in registers.
    const double *scnt = (const double *)tr.bs_cntf + 4 * (int64_t)(first + base);
    const int *sent = (const int *)tr.bs_e + 4 * (int64_t)(first + base);
    // Load set 1 of the step at j0 (nodes past the block repeat its last node).
    auto rows1 = [&](int j0, int (&wv)[kNodesB], int (&wx)[kNodesB][kWidth]) {
#pragma unroll
      for (int u = 0; u < kNodesB; ++u) {
        const int j = min(j0 + u, jn - 1);
        wv[u] = SD(L, roff(rl(nd.x, j)));
#pragma unroll
        for (int r = 0; r < kWidth; ++r) wx[u][r] = SD(L, roff(rl(sxr[r], j)));
      }
    };
    int wvn[kNodesB], wxn[kNodesB][kWidth];
    rows1(0, wvn, wxn);
    sink_stores<2 * kNodesB + 1>(L);  // a step stores flows and a quotient per node, then node flows
    for (int j0 = 0; j0 < jn; j0 += kNodesB) {
      // This step's rows have arrived once only the stores issued after them
      // are outstanding (as in the forward pass).
      wait_vm_left<2 * kNodesB + 1>();
      int wv[kNodesB], wx[kNodesB][kWidth];
#pragma unroll
      for (int u = 0; u < kNodesB; ++u) {
        wv[u] = wvn[u];
#pragma unroll
        for (int r = 0; r < kWidth; ++r) wx[u][r] = wxn[u][r];
      }
      // Load set 2: raw slot words only (counts of nodes whose words hold
      // none, and the tight entries' slots); the arithmetic follows once all
      // are issued, and the phase order is kept as written (see the forward
      // pass).
      __builtin_amdgcn_sched_barrier(0);
      uint64_t svr[kNodesB];
      uint64_t qr[kNodesB][kWidth];
      int tm[kNodesB];
#pragma unroll
      for (int u = 0; u < kNodesB; ++u) {
        const int j = min(j0 + u, jn - 1);
        const uint32_t cv = cnt_of(L, wv[u]);
        svr[u] = slot_or(cv == 0, roff(rl(nd.x, j)), L, cv);  // sigma(v)
        const int dv = dist_of(L, wv[u]);
        const int leafm = rl(nd.w, j);
        tm[u] = 0;
#pragma unroll
        for (int r = 0; r < kWidth; ++r) {
          const int len = rl(slr[r], j);  // kNoLen in unused places
          // Padding lanes (sources past the batch) count as never tight, so
          // their flows are 0 and the lane sums below need no mask.
          const bool t = L.valid & (dv + len == dist_of(L, wx[u][r]));
          tm[u] |= (int)t << r;
          // Q(x) where x has descendants, else sigma(x) (from the word or the slot).
          qr[u][r] = slot_or(t & slot_needed(L, (leafm >> r) & 1, wx[u][r]), roff(rl(sxr[r], j)), L,
                             cnt_of(L, wx[u][r]));
        }
      }
      rows1(j0 + kNodesB, wvn, wxn);
      __builtin_amdgcn_sched_barrier(0);
      double svu[kNodesB];
      double qx[kNodesB][kWidth];  // Q(x) in the lanes where entry r is tight
      // An opaque zero defined here keeps the conversion of the node's count
      // below this point: the compiler would otherwise move it into the
      // count's masked-load branch, where it waits for every outstanding load.
      int zero = 0;
      __asm__ volatile("" : "+s"(zero));
#pragma unroll
      for (int u = 0; u < kNodesB; ++u) {
        const int j = min(j0 + u, jn - 1);
        const int leafm = rl(nd.w, j);
        svu[u] = (double)(svr[u] | (uint32_t)zero);
#pragma unroll
        for (int r = 0; r < kWidth; ++r) {
          // Q(x) where the entry is tight: a quotient from x's backward step,
          // or, where x has no descendants (a uniform bit), 1 / sigma(x).
          double q = __builtin_bit_cast(double, qr[u][r]);
          if ((leafm >> r) & 1)  // the division only where some tight lane's count is not 1
            q = __builtin_amdgcn_ballot_w64(((tm[u] >> r) & 1) & (qr[u][r] != 1))
                    ? 1.0 / (double)qr[u][r]
                    : 1.0;
          qx[u][r] = ((tm[u] >> r) & 1) ? q : 0.0;
        }
      }
      double ns_in[kNodesB];
      int fin_mask = 0;  // the step's pieces that end their node
#pragma unroll
      for (int u = 0; u < kNodesB; ++u) {
        const int j = min(j0 + u, jn - 1);
        const bool live = j0 + u < jn;  // false for the clamped repeats of the last piece
        const int v = rl(nd.x, j), nb = rl(nd.y, j), fl = rl(nd.w, j);  // nb <= kWidth
        const int pos = base + j0 + u;  // result boundaries as in the forward pass
        const bool pfirst = ((fl >> 4) & 1) || (runs && pos == 0);
        const bool fin = live && (((fl >> 5) & 1) || (runs && pos == mine - 1));
        const double svv = svu[u];
        double f[kWidth], acc = 0.0;
#pragma unroll
        for (int r = 0; r < kWidth; ++r) {
          f[r] = ((fl >> (8 + r)) & 1) ? svv * qx[u][r] : cld(scnt, 4 * j + r) * svv * qx[u][r];
          acc += f[r];
        }
        const double fs = wave_sum4(f[0], f[1], kWidth > 2 ? f[2 % kWidth] : 0.0,
                                    kWidth > 3 ? f[3 % kWidth] : 0.0);
        const int e = row_pick(L.lane, cld(sent, 4 * j), cld(sent, 4 * j + 1), cld(sent, 4 * j + 2),
                               cld(sent, 4 * j + 3));
        put_flow_all(L.phi, live & ((L.lane & 15) == 0) & ((L.lane >> 4) < nb), e, fs);
        // The node's dependency part, summed piece by piece in entry order.
        const double tot = pfirst ? acc : carry + acc;
        if (live) carry = tot;
        if constexpr (runs) {
          if (fin) {  // uniform
            sg.x[seg][L.lane] = __builtin_bit_cast(uint64_t, tot);
            sg.y[seg][L.lane] = __builtin_bit_cast(uint64_t, svv);
            if (L.lane == 0) sg.node[seg] = v;
            ++seg;
          }
          put_q_all(L, false, 0, 0.0);
          __asm__ volatile("" ::: "memory");  // not a dead store: the waits count it
          ns_in[u] = 0.0;
        } else {
          double q = 0.0;
          if (fin)  // uniform branch: only the last piece divides, where some count is not 1
            q = __builtin_amdgcn_ballot_w64(svv != 1.0) ? (1.0 + tot) / svv : 1.0 + tot;
          put_q_all(L, fin, roff(v), q);
          ns_in[u] = (fin & L.valid & (L.s != v)) ? tot : 0.0;
          fin_mask |= (int)fin << u;
        }
      }
      double ns = 0.0;  // node flows, summed only in steps that end a node (uniform)
      if (fin_mask)
        ns = wave_sum4(ns_in[0], kNodesB > 1 ? ns_in[1 % kNodesB] : 0.0,
                       kNodesB > 2 ? ns_in[2 % kNodesB] : 0.0, kNodesB > 3 ? ns_in[3 % kNodesB] : 0.0);
      // Node flows at the slot of the node's last piece (consecutive steps,
      // consecutive slots).
      const int u = L.lane >> 4, p0 = first + base;
      const int ps = row_pick(L.lane, p0 + min(j0, jn - 1), p0 + min(j0 + 1, jn - 1),
                              p0 + min(j0 + 2, jn - 1), p0 + min(j0 + 3, jn - 1));
      put_flow_all(L.nodef, ((L.lane & 15) == 0) & (u < kNodesB) & ((fin_mask >> u) & 1), ps, ns);
      wait_vm_left<kNodesB * (1 + kWidth) + 2 * kNodesB + 1>();
    }
  }
}

// The whole sweep of tile t by one workgroup, in state slot `slot`: forward
// over depth levels, backward over depth levels, then the ascending states.
// Every level runs the pipelined steps over its staged pieces; a level split
// in runs (few nodes, many pieces) then merges each node's per-wavefront
// partial results from LDS in a fixed order, so results are deterministic.
// Every record a tile reads it has written itself, so the result does not
// depend on the slot or on what an earlier tile left there.
// With fwd_only, the tile stops after the forward pass (distances and path
// counts left in its slot, for probes of that pass).
template <int kNodesF, int kNodesB>
__device__ __forceinline__ void sweep_tile(const DevIndex &ix, const DevTree &tr, const SweepWork &sw,
                                           int t, int slot, int nsrc, uint64_t *stamps, int fwd_only,
                                           const Segs &sg, double (*sh_acc)[kWave], int wave, int &ovf) {
  uint64_t(*seg_x)[kWave] = sg.x;
  uint64_t(*seg_y)[kWave] = sg.y;
  int *seg_node = sg.node;
  const Lane L = lane_setup(ix, tr, sw, t, slot, nsrc, tr.n);
  const int H = tr.height;
  // Optional phase timestamps (constant 100 MHz clock), one row per tile.
  if (stamps && threadIdx.x == 0) stamps[4 * t] = __builtin_amdgcn_s_memrealtime();

  // ---- forward: distance and path count ----
  for (int k = 0; k < H; ++k) {
    const int nseg = cld(tr.lvl_nseg, k);
    int first, mine;
    level_block(tr.st_off, tr.st_cnt, k, wave, first, mine);
    if (nseg) {
      fwd_level_wave<kNodesF, true, kStage>(ix, tr, L, k, first, mine,
                                            cld(tr.st_seg, k * kSweepWaves + wave), ovf, sg);
    } else {
      const int nar = cld(tr.st_nar, k * kSweepWaves + wave), n3 = cld(tr.st_n3, k * kSweepWaves + wave);
      fwd_level_wave<kNodesF, false, kStage>(ix, tr, L, k, first, mine - nar - n3, 0, ovf, sg);
      fwd_level_wave<kNodesF, false, 3>(ix, tr, L, k, first + mine - nar - n3, n3, 0, ovf, sg);
      fwd_level_wave<kNodesF, false, kNarrowEntries>(ix, tr, L, k, first + mine - nar, nar, 0,
                                                     ovf, sg);
    }
    __syncthreads();
    if (nseg) {
      // Each node's segments, merged in order (min, then saturating count).
      for (int s = wave; s < nseg; s += kSweepWaves) {
        const int v = __builtin_amdgcn_readfirstlane(seg_node[s]);
        if (s > 0 && __builtin_amdgcn_readfirstlane(seg_node[s - 1]) == v) continue;
        int b = (int)seg_y[s][L.lane];
        uint64_t c = seg_x[s][L.lane];
        for (int s2 = s + 1; s2 < nseg && __builtin_amdgcn_readfirstlane(seg_node[s2]) == v; ++s2) {
          const int bt = (int)seg_y[s2][L.lane];
          const uint64_t ct = seg_x[s2][L.lane];
          c = bt < b ? ct : bt == b ? add_sat(c, ct) : c;
          b = min(b, bt);
        }
        put_dist_count(L, roff(v), b, c);
        ovf |= L.valid && c == ~0ull;
      }
      __syncthreads();
    }
  }

  if (stamps && threadIdx.x == 0) stamps[4 * t + 1] = __builtin_amdgcn_s_memrealtime();
  if (fwd_only) return;  // uniform

  // ---- backward: dependencies, shortcut flows, node flows ----
  for (int k = H - 1; k >= 0; --k) {
    const int nseg = cld(tr.blvl_nseg, k);
    int first, mine;
    level_block(tr.bst_off, tr.bst_cnt, k, wave, first, mine);
    if (nseg) {
      bwd_level_wave<kNodesB, true, kStage>(tr, sw, L, first, mine,
                                            cld(tr.bst_seg, k * kSweepWaves + wave), sg);
    } else {
      const int nar = cld(tr.bst_nar, k * kSweepWaves + wave), n3 = cld(tr.bst_n3, k * kSweepWaves + wave);
      bwd_level_wave<kNodesB, false, kStage>(tr, sw, L, first, mine - nar - n3, 0, sg);
      bwd_level_wave<kNodesB, false, 3>(tr, sw, L, first + mine - nar - n3, n3, 0, sg);
      bwd_level_wave<kNodesB, false, kNarrowEntries>(tr, sw, L, first + mine - nar, nar, 0, sg);
    }
    __syncthreads();
    if (nseg) {
      // Each node's segments, summed in order; then its quotient and flow.
      for (int s = wave; s < nseg; s += kSweepWaves) {
        const int v = __builtin_amdgcn_readfirstlane(seg_node[s]);
        if (s > 0 && __builtin_amdgcn_readfirstlane(seg_node[s - 1]) == v) continue;
        double tot = __builtin_bit_cast(double, seg_x[s][L.lane]);
        const double svv = __builtin_bit_cast(double, seg_y[s][L.lane]);
        for (int s2 = s + 1; s2 < nseg && __builtin_amdgcn_readfirstlane(seg_node[s2]) == v; ++s2)
          tot += __builtin_bit_cast(double, seg_x[s2][L.lane]);
        setQ(L, roff(v), (1.0 + tot) / svv);
        const double ns = wave_sum_dpp((L.valid && L.s != v) ? tot : 0.0);
        if (L.lane == 0) sw.node_t[(int64_t)L.t * tr.bslots + cld(tr.bnode_slot, v)] = ns;
      }
      __syncthreads();
    }
  }

  if (stamps && threadIdx.x == 0) stamps[4 * t + 2] = __builtin_amdgcn_s_memrealtime();

  // ---- ascending states: the union of the tile's root paths in top-down
  // order, each node's bag entries split over all wavefronts and merged in
  // wavefront order.  With sigma_A(x) the up-label count where the up-label
  // distance is tight (else 0), the ascending dependency is
  //   dA(x) = sigma_A(x) * (1 + dD(x)) / sigma(x)
  //         + sum over tight bag entries (x -> y) c * sigma_A(x) * dA(y) / sigma_A(y),
  // and x's dependency slot, which holds Q(x) = (1 + dD(x)) / sigma(x) after
  // the backward sweep, is replaced by the quotient
  //   QA(x) = dA(x) / sigma_A(x) = Q(x) + sum over tight entries c * QA(y)
  // (0 where sigma_A(x) = 0).  Entry (x -> y) carries sigma_A(x) * c * QA(y)
  // and x's node flow gains sigma_A(x) times the sum.  Later nodes read only
  // their ancestors' quotients, and a lane whose source is not below x never
  // reads x again.
  {
    const int c0 = cld(sw.cand_off, sw.tile0 + L.t), c1 = cld(sw.cand_off, sw.tile0 + L.t + 1);
    for (int kk = c0; kk < c1; ++kk) {
      const int x = cld(sw.cand, kk);
      const bool anc = x <= L.s && L.s < x + cld(tr.subtree, x);
      const uint32_t ax = roff(x);
      const int wx = SD(L, ax);
      const int dx = dist_of(L, wx);
      uint64_t sa = 0;
      if (anc) {
        const int64_t o = L.rs + cld(ix.depth, x);
        if (ix.dis[o] == dx) sa = ix.cnt[o];
      }
      const double sax = (double)sa;
      double up = 0.0;
      if (__any(sa != 0)) {
        const int e1 = cld(tr.bag_off, x + 1);
        for (int e = cld(tr.bag_off, x) + wave; e < e1; e += 4 * kSweepWaves) {
          int dy[4];
          uint32_t ay[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int eq = e + q * kSweepWaves;
            const int ec = eq < e1 ? eq : e;
            ay[q] = roff(cld(tr.bag_v, ec));
            dy[q] = dist_of(L, SD(L, ay[q]));
          }
          uint64_t qa[4];
          bool t[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int eq = e + q * kSweepWaves;
            const int ec = eq < e1 ? eq : e;
            const int len = eq < e1 ? cld(tr.bag_len, ec) : kNoLen;
            t[q] = sa != 0 && dx + len == dy[q];
            qa[q] = 0;
            if (t[q]) qa[q] = SS(L, ay[q]);  // QA(y): y has descendants
          }
          double g[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int eq = e + q * kSweepWaves;
            const int ec = eq < e1 ? eq : e;
            g[q] = t[q] ? cld(tr.bag_cntf, ec) * __builtin_bit_cast(double, qa[q]) : 0.0;  // c * QA(y)
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) up += g[q];
          const double fs = wave_sum4(L.valid ? sax * g[0] : 0.0, L.valid ? sax * g[1] : 0.0,
                                      L.valid ? sax * g[2] : 0.0, L.valid ? sax * g[3] : 0.0);
          const int eq = e + (L.lane >> 4) * kSweepWaves;
          if ((L.lane & 15) == 0 && eq < e1)
            sw.phi_t[(int64_t)L.t * tr.bags + cld(tr.ibag_pos, eq)] += fs;
        }
      }
      sh_acc[wave][L.lane] = up;
      __syncthreads();
      if (wave == 0) {
        double u = 0.0;
        for (int w = 0; w < kSweepWaves; ++w) u += sh_acc[w][L.lane];
        const bool xleaf = cld(tr.ibag_off, x) == cld(tr.ibag_off, x + 1);
        uint64_t raw = 0;
        if (slot_needed(L, xleaf, wx)) raw = SS(L, ax);
        const double qx = q_from_slot(L, xleaf, wx, raw);
        setQ(L, ax, sa != 0 ? qx + u : 0.0);
        const double ns = wave_sum_dpp((L.valid && x != L.s) ? sax * u : 0.0);
        const int xs = cld(tr.bnode_slot, x);  // -1 without descendants, where ns is 0
        if (L.lane == 0 && xs >= 0) sw.node_t[(int64_t)L.t * tr.bslots + xs] += ns;
      }
      __syncthreads();
    }
  }
  if (stamps && threadIdx.x == 0) stamps[4 * t + 3] = __builtin_amdgcn_s_memrealtime();
}

// State slot of a workgroup's tile.  A launch of no more tiles than slots
// gives workgroup w slot w.  A longer launch (sw.pool > 0) draws on the slots
// as a pool: the hardware keeps starting workgroups as others finish, and
// each takes a free slot with an atomic flag and returns it at its end, so
// the slots, not the launch's tiles, bound the state memory.  The slots are
// split by accelerator die (XCC_ID, on a device of kSweepDies dies), sw.pool
// per die: the dies' L2 caches are not kept coherent with each other inside
// a kernel, so a slot's records never move between dies within a launch.  A die's share holds as many
// slots as workgroups can be resident on it, so a free one appears as soon
// as the workgroup that held it has ended (its return may land a moment after
// the next workgroup starts on the unit, hence the retries).  The wait is
// bounded, kSlotRounds rounds of at most sw.pool attempts and a short sleep
// (about a second), and a workgroup that finds none returns -1 (the launch
// then reports an error).
// s_getreg operand for the die index: hardware register 20 (XCC_ID on gfx950),
// bits [3:0] (offset 0 at bit 6, size 4 encoded as size - 1 at bit 11).
constexpr int kHwRegXccId = 20 | (0 << 6) | ((4 - 1) << 11);
constexpr int kSlotRounds = 1 << 14;
__device__ __forceinline__ int take_slot(const SweepWork &sw) {
  const int die = sw.dies == 1 ? 0 : (int)(__builtin_amdgcn_s_getreg(kHwRegXccId) & 7u);
  const int base = die * sw.pool;
  const int first = (int)(blockIdx.x % (unsigned)sw.pool);
  for (int round = 0; round < kSlotRounds; ++round) {
    for (int i = 0; i < sw.pool; ++i) {
      const int s = base + (first + i) % sw.pool;
      if (atomicCAS(&sw.slot_busy[s], 0, 1) == 0) return s;
    }
    __builtin_amdgcn_s_sleep(64);
  }
  return -1;
}

// One workgroup per tile of 64 sources (tile blockIdx.x of the launch).
template <int kNodesF, int kNodesB, int kWavesPerSimd>
__global__ __launch_bounds__(kSweepThreads, kWavesPerSimd) void sweep_tile_kernel(DevIndex ix, DevTree tr,
                                                                   SweepWork sw, int nsrc,
                                                                   int *ovf_flag,
                                                                   uint64_t *stamps,
                                                                   int fwd_only) {
  __shared__ double sh_acc[kSweepWaves][kWave];
  // Segments of levels split in runs: per-node partial results.
  __shared__ uint64_t seg_x[kSweepMaxSegments][kWave], seg_y[kSweepMaxSegments][kWave];
  __shared__ int seg_node[kSweepMaxSegments];
  __shared__ int sh_slot;
  const Segs sg{seg_x, seg_y, seg_node};
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int t = blockIdx.x;
  if (threadIdx.x == 0) sh_slot = sw.pool > 0 ? take_slot(sw) : t;
  __syncthreads();
  const int slot = __builtin_amdgcn_readfirstlane(sh_slot);
  if (slot < 0) {  // uniform: no slot came free (the host reports it)
    if (threadIdx.x == 0) atomicOr(ovf_flag, kSweepNoSlot);
    return;
  }
  if (threadIdx.x == 0) sw.tile_slot[t] = slot;
  int ovf = 0;
  sweep_tile<kNodesF, kNodesB>(ix, tr, sw, t, slot, nsrc, stamps, fwd_only, sg, sh_acc, wave, ovf);
  if (ovf) atomicOr(ovf_flag, 1);
  if (sw.pool > 0) {
    // Every store of every wave to the slot has reached the die's L2 (a
    // store's count drops when the L2 takes it) before the slot is
    // returned there: the next workgroup to take it, on the same die, writes
    // the same records and reads only what it wrote itself.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) atomicExch(&sw.slot_busy[slot], 0);
  }
}

__global__ void sweep_pool_reset_kernel(int *busy, int slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < slots) busy[i] = 0;
}

using TileKernel = void (*)(DevIndex, DevTree, SweepWork, int, int *, uint64_t *, int);

// Two forms of the tile kernel.  Up to one workgroup per compute unit: 2
// pieces per step in each pass, registers for 5 waves per SIMD (one 16-wave
// workgroup per unit).  More workgroups than units: 1 piece per step,
// registers capped at 64 for 8 waves per SIMD, so two workgroups share each
// unit and one's waits are covered by the other's work (one such workgroup
// alone is 8% slower than the first form; two sharing a unit run about 1.9
// times the tiles per second of one alone).  `groups` is the launch's
// resident workgroups (state slots in use).  The tuning may force either form.
int tile_form(int groups, int cus, int form) {
  if (form == BCX_SWEEP_FORM_ONE_PER_CU || form == BCX_SWEEP_FORM_TWO_PER_CU) return form;
  return groups > cus ? BCX_SWEEP_FORM_TWO_PER_CU : BCX_SWEEP_FORM_ONE_PER_CU;
}
TileKernel tile_kernel(int form) {
  if (form == BCX_SWEEP_FORM_TWO_PER_CU) return (TileKernel)sweep_tile_kernel<1, 1, 8>;
  return (TileKernel)sweep_tile_kernel<2, 2, 1>;
}

// Distance and count rows of a launch's tiles after a forward-only launch:
// row r (tile r / 64, lane r % 64, in the slot that ran the tile) of node p
// goes to d[r * n + p], s[r * n + p] (tree ids), for r below nrows.
__global__ __launch_bounds__(256) void sweep_extract_kernel(SweepWork sw, int n, int dbits, int *d,
                                                            uint64_t *s) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int r = blockIdx.y;
  const char *rec = sw.state + ((int64_t)sw.tile_slot[r / kWave] * (n + 1) + p) * kRec;
  const int lane = r % kWave;
  const int w = *(const int *)(rec + lane * 4);
  const uint32_t cw = (uint32_t)w >> dbits;  // the count, or 0 where it is in the slot
  const uint64_t c = cw ? cw : *(const uint64_t *)(rec + kSlotOff + lane * 8);
  d[(int64_t)r * n + p] = w & (int)((1u << dbits) - 1u);
  s[(int64_t)r * n + p] = c;
}

// node_out[v] (+)= sum over tiles of node_t, phi_out[e] (+)= sum over tiles
// of phi_t, tiles in increasing order.  Both are read in slot order (node
// flows at the slot of the node's last backward piece, shortcut flows in
// inverse-bag order) and each sum written to its node or entry.  Nodes
// without descendants have no node flow: 0, or left as they are when
// accumulating.
__global__ __launch_bounds__(256) void sweep_tile_reduce_kernel(int n, int64_t B, int tiles,
                                                                SweepWork sw, DevTree tr,
                                                                double *node_out, double *phi_out,
                                                                int accumulate) {
  const int64_t S = tr.bslots;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S) {
    const int v = tr.bslot_node[i];
    if (v < 0) return;
    double acc = accumulate ? node_out[v] : 0.0;
    for (int t = 0; t < tiles; ++t) acc += sw.node_t[(int64_t)t * S + i];
    node_out[v] = acc;
  } else if (i < S + n) {
    const int v = (int)(i - S);
    if (!accumulate && tr.bnode_slot[v] < 0) node_out[v] = 0.0;
  } else if (i < S + n + B) {
    const int64_t k = i - S - n;
    const int e = tr.ibag_e[k];
    double acc = accumulate ? phi_out[e] : 0.0;
    for (int t = 0; t < tiles; ++t) acc += sw.phi_t[(int64_t)t * B + k];
    phi_out[e] = acc;
  }
}

// sw.node / sw.phi = the sum over nb per-batch slices, in slice order.
__global__ __launch_bounds__(256) void sweep_sum_batches_kernel(int n, int64_t B, int nb,
                                                                const double *nodeb,
                                                                const double *phib, SweepWork sw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += nodeb[(int64_t)b * n + i];
    sw.node[i] = acc;
  } else if (i < n + B) {
    const int64_t e = i - n;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += phib[(int64_t)b * B + e];
    sw.phi[e] = acc;
  }
}

// Unpacking of the shortcut flows, top-down over the depth levels.  For node
// z and each tight pair (a, b) of its bag (a path a-z-b as short as shortcut
// (a, b)), z carries the share c(a,z) c(z,b) / c(a,b) of that shortcut's flow,
// and the same amount moves onto z's shortcuts (a, z) and (z, b) for deeper
// levels: the restricted-path counting of Reduce (TreeDecomp.cpp:157-193),
// reversed.  Per level, step 1 computes the share of every tight pair of the
// level, step 2 gives each node of the level its interior flow and adds to
// its entries' flows, each sum over pairs in increasing pair order.
//
// Wide levels (at least the tuning's unpack_wide nodes, 1024 by default)
// take two grid-wide launches, one
// thread per pair and one wavefront per node.  Runs of narrower levels go to
// one resident workgroup that walks them with a barrier after each step, so
// the long chain of small top levels costs no launches; there a node gets a
// thread on levels of at least kWave nodes and a whole wavefront otherwise
// (the large bags near the root carry most pairs).  A wavefront's lanes take
// the node's pairs and entries in strides and a fixed lane tree sums them.
// Every sum has a fixed order, so results are deterministic.
constexpr int kUnpackThreads = 1024;

// Sum of pf[idx[p]] for p = p0, p0 + step, ... below p1, added in that
// order; four index loads, then four value loads, are in flight at a time
// (a loop of single loads waits twice per term).
__device__ __forceinline__ double gather_sum(const double *pf, const int *idx, int p, int p1,
                                             int step) {
  double acc = 0.0;
  for (; p + 3 * step < p1; p += 4 * step) {
    const int i0 = idx[p], i1 = idx[p + step], i2 = idx[p + 2 * step], i3 = idx[p + 3 * step];
    const double x0 = pf[i0], x1 = pf[i1], x2 = pf[i2], x3 = pf[i3];
    acc += x0;
    acc += x1;
    acc += x2;
    acc += x3;
  }
  for (; p < p1; p += step) acc += pf[idx[p]];
  return acc;
}

// Pair shares pf[q] for q, q + step, ... below q1 (two at a time): the
// shortcut flow times the pair's fraction of its restricted path count.
__device__ __forceinline__ void pair_flows(const DevTree &tr, const SweepWork &sw, int q, int q1,
                                           int step) {
  for (; q + step < q1; q += 2 * step) {
    const int a0 = tr.tp_ab[q], a1 = tr.tp_ab[q + step];
    const double f0 = tr.tp_frac[q], f1 = tr.tp_frac[q + step];
    const double p0 = sw.phi[a0], p1 = sw.phi[a1];
    sw.pf[q] = p0 * f0;
    sw.pf[q + step] = p1 * f1;
  }
  if (q < q1) sw.pf[q] = sw.phi[tr.tp_ab[q]] * tr.tp_frac[q];
}

__device__ __forceinline__ void unpack_node_serial(const DevTree &tr, const SweepWork &sw, int z) {
  sw.interior[z] = gather_sum(sw.pf, tr.zp_pair, tr.zp_off[z], tr.zp_off[z + 1], 1);
  for (int e = tr.bag_off[z]; e < tr.bag_off[z + 1]; ++e) {
    const int a0 = tr.ep_off[e], a1 = tr.ep_off[e + 1];
    if (a1 > a0) sw.phi[e] += gather_sum(sw.pf, tr.ep_pair, a0, a1, 1);
  }
}

__global__ __launch_bounds__(256) void sweep_pair_flow_kernel(DevTree tr, SweepWork sw, int k) {
  const int q = tr.lvl_pair_off[k] + blockIdx.x * blockDim.x + threadIdx.x;
  if (q < tr.lvl_pair_off[k + 1]) sw.pf[q] = sw.phi[tr.tp_ab[q]] * tr.tp_frac[q];
}

// One node by a whole wavefront (all 64 lanes must come here together).
__device__ __forceinline__ void unpack_node_wave(const DevTree &tr, const SweepWork &sw, int z,
                                                 int lane) {
  const double acc =
      wave_sum_dpp(gather_sum(sw.pf, tr.zp_pair, tr.zp_off[z] + lane, tr.zp_off[z + 1], kWave));
  if (lane == 0) sw.interior[z] = acc;
  for (int e = tr.bag_off[z] + lane; e < tr.bag_off[z + 1]; e += kWave) {
    const int a0 = tr.ep_off[e], a1 = tr.ep_off[e + 1];
    if (a1 > a0) sw.phi[e] += gather_sum(sw.pf, tr.ep_pair, a0, a1, 1);
  }
}

// Four nodes of wide level k per workgroup, one per wavefront.
__global__ __launch_bounds__(256) void sweep_pair_apply_kernel(DevTree tr, SweepWork sw, int k) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int i = tr.lvl_off[k] + blockIdx.x * 4 + wave;
  if (i < tr.lvl_off[k + 1]) unpack_node_wave(tr, sw, __builtin_amdgcn_readfirstlane(tr.lvl_v[i]), lane);
}

__global__ __launch_bounds__(kUnpackThreads) void sweep_unpack_kernel(DevTree tr, SweepWork sw,
                                                                      int k0, int k1) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  constexpr int kWaves = kUnpackThreads / kWave;
  for (int k = k0; k < k1; ++k) {
    pair_flows(tr, sw, tr.lvl_pair_off[k] + tid, tr.lvl_pair_off[k + 1], kUnpackThreads);
    __syncthreads();
    const int lo = tr.lvl_off[k], hi = tr.lvl_off[k + 1];
    if (hi - lo >= kWave) {
      for (int i = lo + tid; i < hi; i += kUnpackThreads) unpack_node_serial(tr, sw, tr.lvl_v[i]);
    } else {
      for (int i = lo + wave; i < hi; i += kWaves)
        unpack_node_wave(tr, sw, __builtin_amdgcn_readfirstlane(tr.lvl_v[i]), lane);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void sweep_combine_kernel(int n, SweepWork sw, double *bc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) bc[v] += 0.5 * (sw.node[v] + sw.interior[v]);
}

__global__ __launch_bounds__(256) void zero_f64_kernel(double *p, int64_t cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = 0.0;
}

inline unsigned blocks_for(int64_t items, int per) { return (unsigned)((items + per - 1) / per); }



}  // namespace

void bcx_ctx::free_sweep() {
  if (sweep_exec) (void)hipGraphExecDestroy(sweep_exec);
  sweep_exec = nullptr;
  sweep_exec_key.clear();
  void *bufs[] = {sw.state, sw.node_t,   sw.phi_t,    sw.node,      sw.phi,
                  sw.interior, sw.tile_src, sw.cand, sw.cand_off, sw.pf, sw.slot_busy, sw.tile_slot};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  sw = SweepWork{};
  plan_pre.clear();
  sweep_tiles = 0;
  sweep_alloc = 0;
  sweep_slots = 0;
  sweep_pool = 0;
}

void bcx_ctx::ensure_sweep(int slots, int tiles) {
  if (slots <= sweep_alloc && tiles <= sweep_tiles) return;
  free_sweep();
  const size_t n = (size_t)ix.n, B = (size_t)tree.bags, T = (size_t)tiles;
  HIP_OK(hipMalloc(&sw.state, (size_t)slots * (n + 1) * kRec));
  HIP_OK(hipMalloc(&sw.node_t, T * (size_t)tree.bslots * 8));
  HIP_OK(hipMalloc(&sw.phi_t, T * std::max<size_t>(B, 1) * 8));
  HIP_OK(hipMalloc(&sw.tile_slot, T * 4));
  HIP_OK(hipMalloc(&sw.slot_busy, (size_t)slots * 4));
  HIP_OK(hipMemset(sw.slot_busy, 0, (size_t)slots * 4));
  HIP_OK(hipMalloc(&sw.node, n * 8));
  HIP_OK(hipMalloc(&sw.phi, std::max<size_t>(B, 1) * 8));
  HIP_OK(hipMalloc(&sw.interior, n * 8));
  HIP_OK(hipMalloc(&sw.pf, std::max<size_t>((size_t)tree.pairs, 1) * 8));
  // The job arrays (tile sources, candidate lists) are sized by sweep_plan.
  sw.slots = slots;
  sweep_alloc = slots;
  sweep_tiles = tiles;
}

namespace bcx_dev {

int sweep_tiles_for(bcx_ctx *c, int32_t count) {
  const int n = c->ix.n;
  const int64_t B = c->tree.bags;
  if (((uint64_t)n + 1) * (uint64_t)kRec > (uint64_t)kDrop || (uint64_t)B * 8 > (uint64_t)kDrop ||
      (uint64_t)c->tree.bslots * 8 > (uint64_t)kDrop)
    throw bcx::Error{BCX_EGRAPH, "bag-sweep tile state exceeds its 32-bit offsets (more than 5242879 vertices)"};
  const int total_tiles = (int)(((int64_t)count + kWave - 1) / kWave);
  const size_t want = (size_t)std::max(1, std::min(c->tn.sweep_batch / kWave, std::max(1, total_tiles)));
  const size_t cus = (size_t)std::max(1, c->cus);
  const int forced = c->tn.sweep_form;
  // Device memory: a state slot's records (one per resident workgroup), and
  // per tile of a launch its node and shortcut flow sums.  Both stay within
  // 85% of the device memory not already held by this context's sweep
  // buffers (and within the tuning's memory limit).
  const size_t per_slot = ((size_t)n + 1) * kRec;
  const size_t per_tile = (size_t)c->tree.bslots * 8 + (size_t)B * 8 + 4;
  auto budget_now = [&]() {
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    const size_t held = (size_t)c->sweep_alloc * per_slot + (size_t)c->sweep_tiles * per_tile;
    size_t cap = (size_t)((double)(fr + held) * 0.85);
    if (c->tn.sweep_mem_limit > 0) cap = std::min(cap, (size_t)c->tn.sweep_mem_limit);
    return cap;
  };
  // The launch's shape for a memory budget:
  //  * up to one tile per compute unit (unless the two-workgroup form is
  //    forced), one slot per tile;
  //  * more, a slot for every workgroup the form can keep resident (two per
  //    unit in the two-workgroup form), and the launch's tiles drawing on
  //    them as a pool, as many tiles as the flow sums' memory allows;
  //  * where the two-workgroup form's slots do not fit, the one-workgroup
  //    form's (chosen form permitting);
  //  * where no pool fits, one slot per tile, a whole number of tiles per
  //    unit beyond one per unit.
  size_t slots = 0, tiles = 0;
  bool cut = false;  // the shape is smaller than the memory-free one
  auto shape = [&](size_t cap) {
    tiles = want;
    cut = false;
    const size_t full1 = (size_t)c->occ_one * cus, full2 = (size_t)c->occ_two * cus;
    const bool pool_ok = (c->dies == 1 || c->dies == kSweepDies) && cus % (size_t)c->dies == 0;
    auto pooled = [&](size_t full) {  // slots for every resident workgroup, as many tiles as fit
      if (!pool_ok || full * (per_slot + per_tile) > cap) return false;
      slots = std::min(want, full);
      tiles = std::max(slots, std::min(want, (cap - slots * per_slot) / per_tile));
      // Cut short by memory, a whole number of tiles per slot, so that the
      // launch's last round does not leave slots idle.
      if (tiles < want) {
        tiles -= tiles % slots;
        cut = true;
      }
      return true;
    };
    if (forced != BCX_SWEEP_FORM_TWO_PER_CU && want <= cus) {
      slots = want;
      if (slots * (per_slot + per_tile) <= cap) return;
    } else if (forced == BCX_SWEEP_FORM_ONE_PER_CU) {
      if (pooled(full1)) return;
    } else {
      if (pooled(full2)) return;
      cut = true;
      if (forced == BCX_SWEEP_FORM_AUTO && pooled(full1)) return;
    }
    cut = true;
    slots = std::min(want, cap / (per_slot + per_tile));
    if (slots > cus) slots -= slots % cus;
    tiles = slots;
  };
  size_t cap = budget_now();
  shape(cap);
  // Cut short by memory with the label-scan workspace allocated: release it
  // (it is re-allocated on its next use) and measure again.
  if (cut && c->ws_slots > 0) {
    HIP_OK(hipStreamSynchronize(c->stream));
    c->free_ws();
    cap = budget_now();
    shape(cap);
  }
  if (slots < 1) throw bcx::Error{BCX_ENOMEM, "not enough device memory for one sweep tile"};
  c->ensure_sweep((int)slots, (int)tiles);
  c->sweep_slots = (int)slots;
  c->sweep_pool = tiles > slots ? (int)(slots / (size_t)c->dies) : 0;
  return (int)tiles;
}

int sweep_launch_form(const bcx_ctx *c, int nt) {
  return tile_form(std::min(nt, c->sweep_slots), c->cus, c->tn.sweep_form);
}

void sweep_occupancy(bcx_ctx *c) {
  int one = 0, two = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &one, (const void *)tile_kernel(BCX_SWEEP_FORM_ONE_PER_CU), kSweepThreads, 0));
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &two, (const void *)tile_kernel(BCX_SWEEP_FORM_TWO_PER_CU), kSweepThreads, 0));
  c->occ_one = std::max(1, one);
  c->occ_two = std::max(1, two);
  int dies = 0;
  HIP_OK(hipDeviceGetAttribute(&dies, hipDeviceAttributeNumberOfXccs, c->device));
  c->dies = std::max(1, dies);
}

// Plans a job over the sorted tree ids pre[0 .. cnt): tile t holds sources
// pre[64 t .. 64 t + 63] (the last tile padded with its first source), and
// its candidate list is the union of its sources' root paths with every
// ancestor first (a source's path stops where it meets the previous
// source's).  Uploads both to the job arrays, growing them when needed.  The
// host plan lives in the context until the next plan, which first waits for
// the stream, so the uploads need no wait of their own.  A job over the same
// sources as the last plan (a repeated full BC, say) keeps the job arrays as
// they are: only plans write them.
void sweep_plan(bcx_ctx *c, const int32_t *pre, int64_t cnt) {
  if (c->sw.tile_src && c->sw.cand && c->sw.cand_off && (size_t)cnt == c->plan_pre.size() &&
      (cnt == 0 || std::memcmp(pre, c->plan_pre.data(), (size_t)cnt * 4) == 0))
    return;
  HIP_OK(hipStreamSynchronize(c->stream));
  c->plan_pre.clear();  // the job arrays change below; recorded again at the end
  const int64_t T = (cnt + kWave - 1) / kWave;
  const auto &par = c->h_parent;
  const auto &sub = c->h_subtree;
  auto &tsrc = c->plan_src;
  auto &cand = c->plan_cand;
  auto &coff = c->plan_cand_off;
  tsrc.assign((size_t)T * kWave, 0);
  cand.clear();
  coff.assign((size_t)T + 1, 0);
  std::vector<int32_t> path;
  for (int64_t t = 0; t < T; ++t) {
    const int64_t t0 = t * kWave;
    const int tn = (int)std::min<int64_t>(kWave, cnt - t0);
    for (int j = 0; j < kWave; ++j) tsrc[(size_t)(t * kWave + j)] = pre[(size_t)(t0 + (j < tn ? j : 0))];
    int prev = -1;
    for (int j = 0; j < tn; ++j) {
      const int s = pre[(size_t)(t0 + j)];
      path.clear();
      for (int x = s; x >= 0; x = par[(size_t)x]) {
        if (prev >= 0 && x <= prev && prev < x + sub[(size_t)x]) break;  // on prev's path
        path.push_back(x);
      }
      for (size_t q = path.size(); q-- > 0;) cand.push_back(path[q]);
      prev = s;
    }
    if (cand.size() > 0x7fffffffu) throw bcx::Error{BCX_EINVAL, "sweep job too large"};
    coff[(size_t)t + 1] = (int32_t)cand.size();
  }
  SweepWork &sw = c->sw;
  if (T > sw.src_cap || !sw.tile_src) {
    if (sw.tile_src) HIP_OK(hipFree(sw.tile_src));
    if (sw.cand_off) HIP_OK(hipFree(sw.cand_off));
    sw.tile_src = nullptr;
    sw.cand_off = nullptr;
    sw.src_cap = 0;
    const int64_t cap = std::max<int64_t>(T, 1);
    HIP_OK(hipMalloc(&sw.tile_src, (size_t)cap * kWave * 4));
    HIP_OK(hipMalloc(&sw.cand_off, (size_t)(cap + 1) * 4));
    sw.src_cap = cap;
  }
  if ((int64_t)cand.size() > sw.cand_cap || !sw.cand) {
    if (sw.cand) HIP_OK(hipFree(sw.cand));
    sw.cand = nullptr;
    sw.cand_cap = 0;
    const int64_t cap = std::max<int64_t>((int64_t)cand.size(), 1);
    HIP_OK(hipMalloc(&sw.cand, (size_t)cap * 4));
    sw.cand_cap = cap;
  }
  if (T > 0)
    HIP_OK(hipMemcpyAsync(sw.tile_src, tsrc.data(), tsrc.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(sw.cand_off, coff.data(), coff.size() * 4, hipMemcpyHostToDevice, c->stream));
  if (!cand.empty())
    HIP_OK(hipMemcpyAsync(sw.cand, cand.data(), cand.size() * 4, hipMemcpyHostToDevice, c->stream));
  c->plan_pre.assign(pre, pre + cnt);
}

// One launch over tiles [tile0, tile0 + nt) of the planned job of nsrc
// sources, in tile-kernel form `form`.  Unless fwd_only, the launch's node
// and shortcut flows then go to node_out[n] / phi_out[B] (added when
// `accumulate`, else stored).  With `timed`, HIP events bracket the kernels
// and per-tile phase stamps are read back (eager launches only).
void sweep_tiles_launch(bcx_ctx *c, int tile0, int nt, int nsrc, int form, bool fwd_only,
                        double *node_out, double *phi_out, bool accumulate, bool timed) {
  const int n = c->ix.n;
  const int64_t B = c->tree.bags;
  if (nt < 1 || nt > c->sweep_tiles) throw bcx::Error{BCX_EINVAL, "sweep launch exceeds its tiles"};
  SweepWork sw = c->sw;
  sw.tile0 = tile0;
  uint64_t *stamps = nullptr;
  if (timed) {
    HIP_OK(hipMalloc(&stamps, (size_t)nt * 4 * 8));
    HIP_OK(hipEventRecord(c->ev[0], c->stream));
  }
  // A launch of more tiles than slots draws on them as a pool; its flags are
  // cleared first by a kernel (not a memset node, so that the launch sequence
  // replays correctly from a captured graph).
  sw.pool = nt > c->sweep_slots ? c->sweep_pool : 0;
  sw.dies = c->dies;
  if (nt > c->sweep_slots && sw.pool < 1) throw bcx::Error{BCX_EINVAL, "sweep launch exceeds its slots"};
  if (sw.pool > 0) {
    sweep_pool_reset_kernel<<<blocks_for(c->sweep_slots, 256), 256, 0, c->stream>>>(sw.slot_busy,
                                                                                  c->sweep_slots);
    HIP_OK(hipGetLastError());
  }
  tile_kernel(form)<<<nt, kSweepThreads, 0, c->stream>>>(c->ix, c->tree, sw, nsrc, c->d_ovf, stamps,
                                                        fwd_only ? 1 : 0);
  HIP_OK(hipGetLastError());
  if (timed) {
    HIP_OK(hipEventRecord(c->ev[1], c->stream));
    HIP_OK(hipEventRecord(c->ev[2], c->stream));
  }
  if (!fwd_only) {
    sweep_tile_reduce_kernel<<<blocks_for(c->tree.bslots + n + B, 256), 256, 0, c->stream>>>(
        n, B, nt, sw, c->tree, node_out, phi_out, accumulate ? 1 : 0);
    HIP_OK(hipGetLastError());
  }
  if (timed) {
    HIP_OK(hipEventRecord(c->ev[3], c->stream));
    HIP_OK(hipEventSynchronize(c->ev[3]));
    float a = 0, b = 0, r = 0;
    HIP_OK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIP_OK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    HIP_OK(hipEventElapsedTime(&r, c->ev[2], c->ev[3]));
    c->timing.scan_ms += a;
    c->timing.dep_ms += b;
    c->timing.reduce_ms += r;
    std::vector<uint64_t> st((size_t)nt * 4);
    HIP_OK(hipMemcpy(st.data(), stamps, st.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(stamps));
    double f = 0, b2 = 0, as = 0;
    for (int t = 0; t < nt; ++t) {
      f += (double)(st[4 * t + 1] - st[4 * t]);
      b2 += (double)(st[4 * t + 2] - st[4 * t + 1]);
      as += (double)(st[4 * t + 3] - st[4 * t + 2]);
    }
    c->timing.sweep_fwd_ms += f / nt * 1e-5;  // 100 MHz ticks
    c->timing.sweep_bwd_ms += fwd_only ? 0.0 : b2 / nt * 1e-5;
    c->timing.sweep_asc_ms += fwd_only ? 0.0 : as / nt * 1e-5;
  }
}

void sweep_launch(bcx_ctx *c, const int32_t *pre, int32_t nsrc, double *node_out,
                  double *phi_out, bool accumulate) {
  const int bt = (nsrc + kWave - 1) / kWave;
  // The launch shape for this many sources (the buffers may have been
  // released for the label-scan workspace since the caller sized its batch).
  if (bt < 1 || bt > sweep_tiles_for(c, nsrc)) throw bcx::Error{BCX_EINVAL, "sweep launch exceeds its tiles"};
  sweep_plan(c, pre, nsrc);
  sweep_tiles_launch(c, 0, bt, nsrc, sweep_launch_form(c, bt), false, node_out, phi_out, accumulate,
                     c->profiling);
  c->timing.scan_launches += 1;
  c->timing.dep_launches += 1;
}

void sweep_sum_batches(bcx_ctx *c, const double *nodeb, const double *phib, int nb) {
  const int n = c->ix.n;
  const int64_t B = c->tree.bags;
  sweep_sum_batches_kernel<<<blocks_for(n + B, 256), 256, 0, c->stream>>>(n, B, nb, nodeb, phib,
                                                                          c->sw);
  HIP_OK(hipGetLastError());
}

void sweep_finish(bcx_ctx *c, double *bc_dev) {
  const int n = c->ix.n;
  const SweepWork sw = c->sw;
  // Unpack: runs of narrow levels in one workgroup, wide levels grid-wide.
  const auto &lvl = c->h_lvl_off;
  const auto &lpo = c->h_lvl_pair_off;
  const int H = c->tree.height;
  const int wide = std::max(1, c->tn.unpack_wide);
  for (int k = 0; k < H;) {
    const int cnt = lvl[(size_t)k + 1] - lvl[(size_t)k];
    if (cnt >= wide) {
      const int qcnt = lpo[(size_t)k + 1] - lpo[(size_t)k];
      if (qcnt > 0)
        sweep_pair_flow_kernel<<<blocks_for(qcnt, 256), 256, 0, c->stream>>>(c->tree, sw, k);
      sweep_pair_apply_kernel<<<blocks_for(cnt, 4), 256, 0, c->stream>>>(c->tree, sw, k);
      ++k;
      continue;
    }
    int k1 = k + 1;
    while (k1 < H && lvl[(size_t)k1 + 1] - lvl[(size_t)k1] < wide) ++k1;
    sweep_unpack_kernel<<<1, kUnpackThreads, 0, c->stream>>>(c->tree, sw, k, k1);
    k = k1;
  }
  sweep_combine_kernel<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, sw, bc_dev);
  HIP_OK(hipGetLastError());
}

void sweep_zero(bcx_ctx *c, double *p, int64_t cnt) {
  if (cnt > 0) zero_f64_kernel<<<blocks_for(cnt, 256), 256, 0, c->stream>>>(p, cnt);
  HIP_OK(hipGetLastError());
}

void run_sweep(bcx_ctx *c, const int32_t *sources, int32_t count, double *bc_inout,
               uint32_t *overflow) {
  const int n = c->ix.n;
  const int64_t B = c->tree.bags;
  std::vector<int32_t> pre((size_t)count);
  for (int32_t k = 0; k < count; ++k) {
    if (sources[k] < 0 || sources[k] >= n) throw bcx::Error{BCX_EINVAL, "source out of range"};
    pre[(size_t)k] = c->iperm[(size_t)sources[k]];
  }
  sort_tree_ids(pre, n);
  const int tiles = sweep_tiles_for(c, count);
  sweep_plan(c, pre.data(), count);
  const int T = (int)(((int64_t)count + kWave - 1) / kWave);
  const int form = sweep_launch_form(c, std::min(tiles, std::max(T, 1)));
  std::vector<double> bc_pre((size_t)n);
  for (int32_t v = 0; v < n; ++v) bc_pre[(size_t)c->iperm[(size_t)v]] = bc_inout[v];
  const int zero = 0;
  HIP_OK(hipMemcpyAsync(c->d_ovf, &zero, 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->d_bc, bc_pre.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  // The job: flows zeroed, one launch (tile kernel + reduction over its
  // tiles) per `tiles` tiles, then the unpack and the BC update.
  auto job = [&](bool timed) {
    sweep_zero(c, c->sw.node, n);
    sweep_zero(c, c->sw.phi, B);
    for (int t0 = 0; t0 < T; t0 += tiles)
      sweep_tiles_launch(c, t0, std::min(tiles, T - t0), count, form, false, c->sw.node, c->sw.phi,
                         true, timed);
    sweep_finish(c, c->d_bc);
  };
  if (c->profiling) {
    job(true);
  } else {
    // Captured once per job shape and buffer set, then replayed: the kernel
    // arguments depend only on these; the sources and candidate lists are
    // data in the job arrays.
    const std::vector<int64_t> key = {count,
                                      tiles,
                                      form,
                                      c->tn.unpack_wide,
                                      c->sweep_slots,
                                      c->sweep_pool,
                                      (int64_t)(intptr_t)c->sw.state,
                                      (int64_t)(intptr_t)c->sw.tile_src,
                                      (int64_t)(intptr_t)c->sw.cand,
                                      (int64_t)(intptr_t)c->sw.cand_off,
                                      (int64_t)(intptr_t)c->d_bc};
    if (!c->sweep_exec || c->sweep_exec_key != key) {
      if (c->sweep_exec) HIP_OK(hipGraphExecDestroy(c->sweep_exec));
      c->sweep_exec = nullptr;
      c->sweep_exec_key.clear();
      hipGraph_t g;
      HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      try {
        job(false);
      } catch (...) {
        hipGraph_t dead;
        (void)hipStreamEndCapture(c->stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      HIP_OK(hipStreamEndCapture(c->stream, &g));
      const hipError_t e = hipGraphInstantiate(&c->sweep_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_OK(e);
      c->sweep_exec_key = key;
    }
    HIP_OK(hipGraphLaunch(c->sweep_exec, c->stream));
  }
  const int launches = (T + tiles - 1) / tiles;
  c->timing.scan_launches += launches;
  c->timing.dep_launches += launches;
  c->timing.sources += count;
  int ovf = 0;
  HIP_OK(hipMemcpyAsync(bc_pre.data(), c->d_bc, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(&ovf, c->d_ovf, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (ovf & kSweepNoSlot) throw bcx::Error{BCX_EDEVICE, "bag sweep: a workgroup found no free state slot"};
  for (int32_t v = 0; v < n; ++v) bc_inout[v] = bc_pre[(size_t)c->iperm[(size_t)v]];
  if (overflow) *overflow = (uint32_t)(ovf & 1);
}

int sweep_query(bcx_ctx *c, const int32_t *sources, int32_t count, int32_t *d, uint64_t *sigma) {
  const int n = c->ix.n;
  if (count <= 0) return 0;
  std::vector<std::pair<int32_t, int32_t>> key((size_t)count);
  for (int32_t k = 0; k < count; ++k) {
    if (sources[k] < 0 || sources[k] >= n) throw bcx::Error{BCX_EINVAL, "source out of range"};
    key[(size_t)k] = {c->iperm[(size_t)sources[k]], k};
  }
  std::stable_sort(key.begin(), key.end());
  std::vector<int32_t> pre((size_t)count);
  for (int32_t k = 0; k < count; ++k) pre[(size_t)k] = key[(size_t)k].first;
  const int tiles = sweep_tiles_for(c, count);
  // Launches of up to `chunk` tiles, their rows extracted to about 1 GiB of
  // scratch at a time (and at most 65535 rows, the grid's second dimension).
  const int64_t rows_cap = std::max<int64_t>(kWave, (1ll << 30) / (12ll * n));
  // No pooled launches here: a tile's records must stay in its slot until
  // they are extracted.
  const int chunk = (int)std::max<int64_t>(
      1, std::min<int64_t>({(int64_t)tiles, (int64_t)c->sweep_slots, rows_cap / kWave, 1023}));
  const int form = sweep_launch_form(c, chunk);
  sweep_plan(c, pre.data(), count);
  const int zero = 0;
  HIP_OK(hipMemcpyAsync(c->d_ovf, &zero, 4, hipMemcpyHostToDevice, c->stream));
  const int T = (count + kWave - 1) / kWave;
  int *dd = nullptr;
  uint64_t *ds = nullptr;
  auto release = [&]() {
    if (dd) (void)hipFree(dd);
    if (ds) (void)hipFree(ds);
  };
  try {
    HIP_OK(hipMalloc(&dd, (size_t)chunk * kWave * n * 4));
    HIP_OK(hipMalloc(&ds, (size_t)chunk * kWave * n * 8));
    std::vector<int32_t> hd;
    std::vector<uint64_t> hs;
    for (int t0 = 0; t0 < T; t0 += chunk) {
      const int nt = std::min(chunk, T - t0);
      const int rows = std::min(nt * kWave, count - t0 * kWave);
      sweep_tiles_launch(c, t0, nt, count, form, true, nullptr, nullptr, false, c->profiling);
      sweep_extract_kernel<<<dim3((unsigned)((n + 255) / 256), (unsigned)rows), 256, 0, c->stream>>>(
          c->sw, n, c->tree.dbits, dd, ds);
      HIP_OK(hipGetLastError());
      hd.resize((size_t)rows * n);
      hs.resize((size_t)rows * n);
      HIP_OK(hipMemcpyAsync(hd.data(), dd, hd.size() * 4, hipMemcpyDeviceToHost, c->stream));
      HIP_OK(hipMemcpyAsync(hs.data(), ds, hs.size() * 8, hipMemcpyDeviceToHost, c->stream));
      HIP_OK(hipStreamSynchronize(c->stream));
      for (int r = 0; r < rows; ++r) {
        const int64_t i = key[(size_t)(t0 * kWave + r)].second;
        for (int32_t v = 0; v < n; ++v) {
          const size_t p = (size_t)r * n + (size_t)c->iperm[(size_t)v];
          if (d) d[i * n + v] = hd[p];
          if (sigma) sigma[i * n + v] = hs[p];
        }
      }
    }
  } catch (...) {
    release();
    throw;
  }
  release();
  int ovf = 0;
  HIP_OK(hipMemcpy(&ovf, c->d_ovf, 4, hipMemcpyDeviceToHost));
  if (ovf & kSweepNoSlot) throw bcx::Error{BCX_EDEVICE, "bag sweep: a workgroup found no free state slot"};
  return (ovf & 1) ? 1 : 0;
}

void sweep_check_slots(bcx_ctx *c) {
  int ovf = 0;
  HIP_OK(hipMemcpy(&ovf, c->d_ovf, 4, hipMemcpyDeviceToHost));
  if (ovf & kSweepNoSlot) throw bcx::Error{BCX_EDEVICE, "bag sweep: a workgroup found no free state slot"};
}

}  // namespace bcx_dev
