/*
 * bcx.h — C ABI of the MI355X-native QBC / MHL-QBC / RQ betweenness path.
 *
 * Plain pointers and sizes only.  Every entry point returns BCX_OK (0) or a
 * negative BCX_E* code; bcx_last_error() returns a thread-local message for
 * the last failing call on the calling thread.  The caller owns every host
 * array it passes; the library copies what it needs and owns device memory
 * per handle.  Vertex ids at this boundary are the 0-based ids of the graph
 * file (file id - 1), as in the reference's Graph class.
 *
 * Reference interfaces replaced (CSU-WRZ/GPU_BetweennessCentrality):
 *   graph reading ............ Graph::ReadGraph (src/Graph.cpp:4-85) and
 *                              Graph_::ReadWeightedGraph (src/TreeCount/Graph.cpp:41-81)
 *   count-index build ........ TreeDecomp ctor/Reduce/MakeTree/MakeIndex
 *                              (src/TreeCount/TreeDecomp.cpp:12-511), called at
 *                              MHLQBC.cu:497-501 and RQ.cu:470-474
 *   order-file index build ... Graph::CHconsorderMT/makeTree/makeIndex
 *                              (src/buildIndex.cpp:4-278), called at QBC.cu:421-423
 *   device upload ............ moveCountIndexToDevice (MHLQBC.cu:6-128),
 *                              moveDisIndexToDevice (QBC.cu:11-127),
 *                              initGraphTravel (MHLQBC.cu:130-160)
 *   full BC .................. Graph::brandesBCIndexGPU_5 (MHLQBC.cu:495-537,
 *                              QBC.cu:419-584) with BC_Kernel/addKernel
 *   subset BC ................ Graph::brandesBCIndexGPU_PointBatch (RQ.cu:465-531)
 *   (d, sigma, hop) probe .... stage 2 of BC_Kernel (MHLQBC.cu:334-412)
 */
#ifndef BCX_H_
#define BCX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BCX_OK = 0,
  BCX_EINVAL = -1,     /* bad argument or malformed input file            */
  BCX_EIO = -2,        /* file cannot be opened / read / written          */
  BCX_EGRAPH = -3,     /* graph violates a precondition (see message)     */
  BCX_ENOMEM = -4,     /* host or device allocation failed                */
  BCX_EDEVICE = -5,    /* HIP runtime error                               */
  BCX_EBUSY = -6,      /* RQ server: request cannot be admitted right now */
  BCX_ENOTFOUND = -7,  /* RQ server: unknown request id                   */
};

/* Execution forms of the full-BC path. */
enum {
  BCX_MODE_PLAIN = 0,   /* one source per launch sequence                 */
  BCX_MODE_BATCHED = 1, /* source tiles, many sources per launch, graphs  */
  BCX_MODE_SWEEP = 2,   /* bag sweep over the tree, 64-source tiles        */
};

typedef struct bcx_graph bcx_graph; /* parsed graph, host memory           */
typedef struct bcx_index bcx_index; /* tree-decomposition label index      */
typedef struct bcx_ctx bcx_ctx;     /* device-resident graph + index       */
typedef struct bcx_rq bcx_rq;       /* RQ serving engine on one context    */

typedef struct bcx_index_stats {
  int32_t n;              /* vertices                                       */
  int32_t tree_height;    /* max depth + 1 (reference tree_height_)         */
  int32_t tree_width;     /* max bag size (reference tree_width_)           */
  int32_t count_overflow; /* 1 if some label count saturated at UINT64_MAX;
                             informational: results are flagged only when
                             their own final count saturates               */
  int64_t label_entries;  /* sum over nodes of (depth + 1)                  */
  double mean_depth;      /* label_entries / n                              */
  double build_seconds;   /* elimination + tree + labels, wall clock        */
  int64_t bag_entries;    /* sum over nodes of the bag size (shortcuts to
                             ancestors): the bag-sweep's work per source    */
} bcx_index_stats;

typedef struct bcx_tuning {
  int32_t tile;          /* sources per label-scan tile (= 64, one per lane) */
  int32_t tchunk;        /* targets per wavefront work item                 */
  int32_t bp_threads;    /* threads per dependency group: 64..512, pow2   */
  int32_t batch;         /* sources per batched launch sequence             */
  int32_t sweep_batch;   /* sources per bag-sweep launch (64 per tile, each
                            tile run by one workgroup, the launch's
                            resident workgroups taking tiles in turn);
                            default 32 x 64 x compute units (two workgroups
                            per unit at a time)                             */
  int32_t sweep_form;    /* bag-sweep tile-kernel form: BCX_SWEEP_FORM_AUTO
                            (by tiles per launch), _ONE_PER_CU (2 pieces per
                            step, one workgroup per compute unit) or
                            _TWO_PER_CU (1 piece per step, registers capped
                            so two workgroups share a unit).  Every form
                            gives bitwise the same results.                */
  int32_t unpack_wide;   /* bag-sweep unpack: depth levels of at least this
                            many nodes run as two grid-wide launches, the
                            rest in one resident workgroup (default 1024)  */
  int64_t sweep_mem_limit; /* bytes of device memory a sweep launch's state
                            may use beyond the 85%-of-free cap (0: no
                            further limit)                                  */
} bcx_tuning;

enum {
  BCX_SWEEP_FORM_AUTO = 0,
  BCX_SWEEP_FORM_ONE_PER_CU = 1,
  BCX_SWEEP_FORM_TWO_PER_CU = 2,
};

typedef struct bcx_timing {
  double scan_ms;     /* label scan (distance, count, hop), bag-sweep tile
                         and H2H distance-scan kernels                     */
  double dep_ms;      /* DAG + layered dependency kernels                  */
  double reduce_ms;   /* BC accumulation kernels                           */
  int64_t scan_launches;
  int64_t dep_launches;
  int64_t sources;
  double sweep_fwd_ms;  /* bag sweep: mean per-tile phase times (profiling) */
  double sweep_bwd_ms;
  double sweep_asc_ms;
} bcx_timing;

const char *bcx_last_error(void);
const char *bcx_version(void);

/* ---- graph ------------------------------------------------------------ */
/* Reads "n m" then m lines "u v w" (1-based ids).  Each undirected edge must
 * be listed in both directions with equal weight w >= 1; duplicate lines are
 * dropped first-wins as in src/Graph.cpp:29-35.  Violations return BCX_EGRAPH
 * (the reference continues on such input; see src/Graph.cpp:7-9). */
int bcx_graph_read(const char *path, bcx_graph **out);
/* Same contract from memory: arcs (u[i], v[i], w[i]), 0-based ids. */
int bcx_graph_from_arcs(int32_t n, int64_t arcs, const int32_t *u,
                        const int32_t *v, const int32_t *w, bcx_graph **out);
int bcx_graph_info(const bcx_graph *g, int32_t *n, int64_t *arcs);
/* 1 if connected (every path precondition of the index needs it). */
int bcx_graph_connected(const bcx_graph *g);
void bcx_graph_free(bcx_graph *g);

/* ---- index ------------------------------------------------------------ */
/* Min-degree elimination index with distance, uint64 path-count and max-hop
 * labels (TreeDecomp::Reduce/MakeTree/MakeIndex, TreeDecomp.cpp:54-511). */
int bcx_index_build_mhl(const bcx_graph *g, bcx_index **out);
/* Same labels, elimination in the order of an order file ("n" then
 * "nodeID order" lines, 0-based ids; Graph::CHconsorderMT, buildIndex.cpp:4-30). */
int bcx_index_build_order(const bcx_graph *g, const char *order_path,
                          bcx_index **out);
/* Same, from an in-memory order: order[v] = contraction rank of vertex v. */
int bcx_index_build_with_order(const bcx_graph *g, const int32_t *order,
                               bcx_index **out);
/* Index build flags. */
enum {
  /* Build the tree and bags on the host and leave the label rows to the
   * device: bcx_ctx_create computes them there, level by level (the label
   * pass of TreeDecomp::MakeIndex, TreeDecomp.cpp:355-477, is the larger
   * part of a host build at a million vertices, and its rows are the largest
   * host array).  Such an index has no host label rows: bcx_index_query,
   * bcx_index_save and the label rows of bcx_index_export return
   * BCX_EINVAL, and its count_overflow covers the bags only (the contexts'
   * results flag their own saturated counts as ever). */
  BCX_INDEX_DEVICE_LABELS = 1,
};
/* bcx_index_build_mhl (order_path and order NULL), bcx_index_build_order
 * (order_path) or bcx_index_build_with_order (order), with flags
 * (BCX_INDEX_DEVICE_LABELS or 0). */
int bcx_index_build_flags(const bcx_graph *g, const char *order_path,
                          const int32_t *order, uint32_t flags,
                          bcx_index **out);
int bcx_index_stats_get(const bcx_index *ix, bcx_index_stats *st);
/* Host-side (d, sigma, hop) query on the labels; used by tools and tests. */
int bcx_index_query(const bcx_index *ix, int32_t s, int32_t t, int32_t *d,
                    uint64_t *sigma, uint32_t *hop);
/* Min-degree contraction order (replaces Graph::GraphConstract,
 * src/buildIndex.cpp:759-925, which writes QBC's order files): order[v] =
 * contraction rank of v (n entries).  shortcut_arcs (may be NULL) receives
 * the number of contracted-neighbour arcs (the reference's "super edge
 * number"); width_log (may be NULL, n / 100000 + 1 entries) the largest
 * contracted neighbourhood before every 100,000th contraction. */
int bcx_min_degree_order(const bcx_graph *g, int32_t *order, int64_t *shortcut_arcs,
                         int32_t *width_log);
/* QBC's H2H distance labels on this index's tree (replaces the global-distance
 * labelling of Graph::makeIndexDFS, src/buildIndex.cpp:222-278, and the
 * separator reduction Graph::optimizeVS, src/buildIndex.cpp:313-386): the
 * shortest distance from every node to each ancestor, and per node the
 * separator scanned when it is the LCA of a pair; and the P2H branch tables
 * of Graph::SVSP_P2H / P2H_3 (src/buildIndex.cpp:565-719) with the
 * reference's k = 10, d = 5, R = 1e6 (QBC.cu:431) and a fixed seed where the
 * reference seeds from the clock (src/buildIndex.cpp:608): for the kept
 * branch nodes, per descendant the separator entries not dominated from it.
 * Needed before bcx_ctx_create for bcx_h2h_distances.  Host memory: 4 bytes
 * per label entry plus the separators and tables. */
int bcx_index_build_h2h(bcx_index *ix);
/* Mean number of label entries the H2H distance query reads over `pairs`
 * seeded random pairs: with the optimizeVS separators only, and with the P2H
 * branch tables as the device scan uses them. */
int bcx_index_h2h_scan_stats(const bcx_index *ix, int64_t pairs, uint64_t seed,
                             double *mean_separator, double *mean_p2h);
/* Host-side H2H distance query (vertex ids); requires bcx_index_build_h2h. */
int bcx_index_h2h_query(const bcx_index *ix, int32_t s, int32_t t, int32_t *d);
/* order[p] = vertex id of the p-th tree node in DFS preorder (n entries).
 * Sources processed in this order share label prefixes best. */
int bcx_index_tree_order(const bcx_index *ix, int32_t *order);
/* Tree shape per vertex v: parent[v] = parent vertex id (-1 for the root),
 * height[v] = depth + 1 (the reference's Node_::height, TreeDecomp.h:106). */
int bcx_index_tree(const bcx_index *ix, int32_t *parent, int32_t *height);
/* Mean scan length pos(lca)+1 over `pairs` seeded random pairs. */
int bcx_index_mean_scan(const bcx_index *ix, int64_t pairs, uint64_t seed,
                        double *mean);
/* Write / read a device-ready binary image of the index (own format; replaces
 * the headerless tmpRoadCountIndex dump, TreeDecomp.cpp:460-467,496-506).  The
 * image holds the tree, labels, bags, H2H labels and separators when built,
 * and a fingerprint of the graph (arc count and a hash of the arcs) that
 * bcx_ctx_create checks.  Loading validates every array (BCX_EIO with the
 * first violation): tree shape and preorder, row lengths, own entries, bag
 * entries being ancestors in depth order ending at the parent, value ranges,
 * and the distance bound. */
int bcx_index_save(const bcx_index *ix, const char *path);
int bcx_index_load(const char *path, bcx_index **out);
/* The host-array seam of moveCountIndexToDevice (MHLQBC.cu:6-128): an index
 * from the reference's TreeDecomp data, indexed by vertex id (0-based).
 *   parent[v]            tree parent vertex, -1 for the root (Node_::parent)
 *   label_off[n + 1]     v's label row is [label_off[v], label_off[v + 1]),
 *                        depth(v) entries over its ancestors, root first
 *   dis / cnt / hop      the rows' restricted distance, shortest-path count
 *                        and max hop count (Node_::dis, cnt, my_cnt,
 *                        TreeDecomp.h:95-118)
 *   bag_off[n + 1]       v's bag is [bag_off[v], bag_off[v + 1])
 *   bag_v / bag_len /    its neighbours at elimination with the shortcut's
 *   bag_cnt / bag_hop    length, count and hop count (Node_::vertices,
 *                        vertices_length, TreeDecomp::edge_count_,
 *                        my_edge_count_, TreeDecomp.cpp:157-200)
 * g (may be NULL) records the graph's fingerprint for bcx_ctx_create.  The
 * arrays are copied and validated as a loaded image is (BCX_EIO or
 * BCX_EINVAL on the first violation). */
int bcx_index_from_arrays(int32_t n, const int32_t *parent,
                          const int64_t *label_off, const int32_t *dis,
                          const uint64_t *cnt, const uint32_t *hop,
                          const int64_t *bag_off, const int32_t *bag_v,
                          const int32_t *bag_len, const uint64_t *bag_cnt,
                          const uint32_t *bag_hop, const bcx_graph *g,
                          bcx_index **out);
/* The same arrays out of an index (each pointer may be NULL): label rows
 * hold label_entries - n entries and the bags bag_entries (bcx_index_stats). */
int bcx_index_export(const bcx_index *ix, int32_t *parent, int64_t *label_off,
                     int32_t *dis, uint64_t *cnt, uint32_t *hop,
                     int64_t *bag_off, int32_t *bag_v, int32_t *bag_len,
                     uint64_t *bag_cnt, uint32_t *bag_hop);
/* The host-array seam of moveDisIndexToDevice (QBC.cu:11-127): QBC's H2H
 * distance index as the reference's Graph holds it after makeIndex,
 * optimizeVS and SVSP_P2H (src/buildIndex.cpp:208-278, 313-386, 565-719;
 * fields of Node, include/Graph.h:16-62), indexed by vertex id (0-based),
 * installed on this index's tree (in place of bcx_index_build_h2h):
 *   parent[v]             tree parent vertex, -1 for the root (Node::pa);
 *                         must equal the index's tree (an index built from
 *                         the same order file has it)
 *   gdis_off[n + 1], gdis v's row [gdis_off[v], gdis_off[v + 1]): depth(v)
 *                         shortest distances to its ancestors, root first
 *                         (Node::dis)
 *   sep_off[n + 1], sep   the row positions scanned when v is the LCA of a
 *                         pair, v's own position included (Node::posOptimize;
 *                         a position is an ancestor depth)
 *   branch[nbranch]       the kept P2H branch vertices (Node::isBranchNode)
 *   p2h_off[nbranch * n + 1], p2h
 *                         for branch[j] and vertex s, the positions
 *                         [p2h_off[j * n + s], p2h_off[j * n + s + 1]) of p2h
 *                         (Node::posP2H_3[s]); empty unless s is below
 *                         branch[j]
 * Every array is checked first (BCX_EINVAL for shape, BCX_EIO for ranges);
 * the index is unchanged by a call that fails.  nbranch may be 0 (branch,
 * p2h_off and p2h may then be NULL). */
int bcx_index_h2h_from_arrays(bcx_index *ix, const int32_t *parent,
                              const int64_t *gdis_off, const int32_t *gdis,
                              const int64_t *sep_off, const int32_t *sep,
                              int32_t nbranch, const int32_t *branch,
                              const int64_t *p2h_off, const int32_t *p2h);
/* Sizes of the H2H arrays of an index with H2H labels: gdis entries
 * (label_entries - n), separator entries, kept branch vertices and P2H list
 * entries (each pointer may be NULL). */
int bcx_index_h2h_sizes(const bcx_index *ix, int64_t *gdis_entries,
                        int64_t *sep_entries, int32_t *nbranch,
                        int64_t *p2h_entries);
/* The same arrays out, in bcx_index_h2h_from_arrays's layout (each pointer
 * may be NULL; p2h_off has nbranch * n + 1 entries). */
int bcx_index_h2h_export(const bcx_index *ix, int64_t *gdis_off, int32_t *gdis,
                         int64_t *sep_off, int32_t *sep, int32_t *branch,
                         int64_t *p2h_off, int32_t *p2h);
void bcx_index_free(bcx_index *ix);

/* ---- device context --------------------------------------------------- */
/* Uploads the index (moveCountIndexToDevice, MHLQBC.cu:6-128), computing the
 * label rows on the device when the index leaves them there
 * (BCX_INDEX_DEVICE_LABELS); refuses an index of another graph. */
int bcx_ctx_create(int device, const bcx_graph *g, const bcx_index *ix,
                   bcx_ctx **out);
/* The context's label rows in bcx_index_export's layout (label_off[n + 1],
 * rows of label_entries - n entries; each pointer may be NULL), copied from
 * the device; for checking rows computed there against a host build.  No
 * reference counterpart. */
int bcx_ctx_export_labels(const bcx_ctx *ctx, int64_t *label_off, int32_t *dis,
                          uint64_t *cnt, uint32_t *hop);
void bcx_ctx_free(bcx_ctx *ctx);
int bcx_ctx_set_tuning(bcx_ctx *ctx, const bcx_tuning *tn);
int bcx_ctx_get_tuning(const bcx_ctx *ctx, bcx_tuning *tn);
/* Autotuner (no reference counterpart: the reference fixes 256x256 launches
 * at compile time, MHLQBC.cu:3).  Times candidate label-scan target chunks
 * and dependency-group widths on `sources` (one batch at most), and bag-sweep
 * launch sizes (2, 4, 8, 16 and 32 tiles of 64 sources per compute unit,
 * those that fit device memory) in both tile-kernel forms, each on the same
 * job of 32 x 64 sources per compute unit taken in tree order (at most n);
 * keeps the fastest of each in
 * the context (the current sweep setting unless a candidate is 2% faster;
 * settings that run that job as the same launches are timed once) and
 * copies the tuning to *best (may be NULL). */
int bcx_ctx_autotune(bcx_ctx *ctx, const int32_t *sources, int32_t count,
                     bcx_tuning *best);
/* Sources one bag-sweep launch takes (the tuning's sweep_batch, cut to whole
 * tiles of 64 by device memory); allocates the launch state. */
int bcx_ctx_sweep_capacity(bcx_ctx *ctx, int32_t *sources_per_launch);
/* Bag-sweep state slots allocated (resident workgroups of a full launch:
 * two per compute unit in the two-workgroup form, each with its own tile
 * records) and the device's compute units; no reference counterpart (the
 * reference gives every source its own block, MHLQBC.cu:508). */
int bcx_ctx_sweep_slots(const bcx_ctx *ctx, int32_t *slots, int32_t *compute_units);
/* Bytes of device memory one batched sequence of `batch` sources needs. */
int bcx_ctx_workspace_bytes(const bcx_ctx *ctx, int32_t batch, size_t *bytes);
int bcx_ctx_mem_info(const bcx_ctx *ctx, size_t *free_bytes, size_t *total_bytes);
int bcx_ctx_timing(const bcx_ctx *ctx, bcx_timing *t);
void bcx_ctx_timing_reset(bcx_ctx *ctx);
/* on != 0: launch batches kernel by kernel with HIP events between phases
 * (accumulated into bcx_timing) instead of replaying a captured graph. */
int bcx_ctx_set_profiling(bcx_ctx *ctx, int on);
/* Measurement hook (no reference counterpart): one device-to-device copy of
 * `bytes`, timed; used to calibrate HBM traffic counters. */
int bcx_ctx_copy_probe(bcx_ctx *ctx, size_t bytes, double *ms);
/* Measurement hook (no reference counterpart): one streaming read of `bytes`
 * (a multiple of 16, 16-byte loads), timed; calibrates the HBM read
 * counters beside the copy. */
int bcx_ctx_read_probe(bcx_ctx *ctx, size_t bytes, double *ms);

/* ---- hot path --------------------------------------------------------- */
/* Full BC over all sources: bc_out[n] (undirected, unnormalised, each
 * unordered pair once; MHLQBC.cu:470 / QBC.cu:395).  overflow (may be NULL)
 * receives 1 if any path count overflowed uint64. */
int bcx_full_bc(bcx_ctx *ctx, int mode, double *bc_out, uint32_t *overflow);
/* Dependencies of a source subset added into bc_inout[n] (0.5 * delta per
 * source).  Sources are accumulated in increasing tree-order, so a result
 * does not depend on the mode or on how the set is split into calls. */
int bcx_bc_sources(bcx_ctx *ctx, int mode, const int32_t *sources,
                   int32_t count, double *bc_inout, uint32_t *overflow);
/* Distance, path count and max-hop layer from `source` to every vertex. */
int bcx_query_dist_count(bcx_ctx *ctx, int32_t source, int32_t *d,
                         uint64_t *sigma, uint32_t *hop);
/* The same rows for `count` sources, as computed by one form of the full-BC
 * path (stage 2 of BC_Kernel, MHLQBC.cu:334-412): BCX_MODE_PLAIN (thread per
 * target), BCX_MODE_BATCHED (the 64-source label-scan tile) or BCX_MODE_SWEEP
 * (the forward pass of the bag-sweep tile kernel, in the kernel form the
 * tuning selects).  Row i of d / sigma / hop (i * n + v) belongs to
 * sources[i].  The bag sweep orders its Brandes pass by tree depth and keeps
 * no hop layers: with BCX_MODE_SWEEP, hop must be NULL.  Returns 1 instead of
 * BCX_OK when some count saturated at UINT64_MAX. */
int bcx_query_dist_count_mode(bcx_ctx *ctx, int mode, const int32_t *sources,
                              int32_t count, int32_t *d, uint64_t *sigma,
                              uint32_t *hop);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same rows as bcx_query_dist_count (all three outputs
 * required), computed by the same kernel, with no allocation of any kind.
 * Its device staging lives in a fixed array of the library (48 MiB per GPU,
 * held until the process ends) shared by every call on that GPU: graphs of
 * more than 1,572,864 vertices are refused with BCX_EINVAL, and calls that
 * overlap on one GPU (other threads, other contexts) race. */
int bcx_query_dist_count_noalloc(bcx_ctx *ctx, int32_t source, int32_t *d,
                                 uint64_t *sigma, uint32_t *hop);
/* QBC distance-only scan (replaces the distance part of BC_Kernel,
 * QBC.cu:243-331: LCA by RMQ, then the ancestor lookup or the minimum over
 * the LCA's separator of the two H2H labels): dist_out[i * n + v] = d(sources[i],
 * v), exact (999999999 when unreachable).  The index must have had
 * bcx_index_build_h2h called before the context was created; the context
 * then holds the labels on the device (4 bytes per entry, interleaved by 64
 * tree-order neighbours, about 1% padding on road graphs).  Scratch per call:
 * two device buffers of min(count, 2^30 / (4 n), 65535) rows. */
int bcx_h2h_distances(bcx_ctx *ctx, const int32_t *sources, int32_t count,
                      int32_t *dist_out);
/* Subset BC (RQ): bc_out[q] = BC of points[i] (RQ.cu:451-459, summed). */
int bcx_subset_bc(bcx_ctx *ctx, const int32_t *points, int32_t q,
                  double *bc_out);

/* ---- RQ serving engine ------------------------------------------------ */
/* Continuous-batching subset-BC server on one context: requests are admitted
 * against a device-memory budget and ride one cyclic source sweep; each
 * completes after a full cycle and its result is summed in a fixed order, so
 * it is bitwise the same alone or co-batched.  Driven by bcx_rq_step().
 * Replaces the reference's one-shot brandesBCIndexGPU_PointBatch
 * (RQ.cu:465-531): the result equals its BC_Kernel_merge sum (RQ.cu:451-459).
 * bcx_rq_create_mode picks the form of the cyclic sweep: BCX_MODE_SWEEP (per
 * batch node and shortcut flows kept in shared slots; a finished request is
 * unpacked once) or BCX_MODE_BATCHED (label scan; per request a [batch][q]
 * table).  bcx_rq_create is bcx_rq_create_mode with BCX_MODE_SWEEP. */
/* hbm_budget_bytes bounds the engine's device memory: its fixed part (in the
 * bag-sweep form one node and one shortcut flow slot per batch, nb (n + B) 8
 * bytes, and n 8 bytes of batch sums; in the label-scan form the source list
 * and batch sums, 12 n bytes) is taken at creation (BCX_EINVAL if the budget
 * is smaller), and each request takes bcx_rq_bytes_needed(q) while in flight
 * and returns it when it completes or is cancelled. */
int bcx_rq_create(bcx_ctx *ctx, size_t hbm_budget_bytes, bcx_rq **out);
int bcx_rq_create_mode(bcx_ctx *ctx, int mode, size_t hbm_budget_bytes,
                       bcx_rq **out);
void bcx_rq_free(bcx_rq *rq);
/* Admit a request; *id receives its handle.  BCX_EBUSY if it does not fit now. */
int bcx_rq_submit(bcx_rq *rq, const int32_t *points, int32_t q, int64_t *id);
int bcx_rq_cancel(bcx_rq *rq, int64_t id);
/* Run one source batch of the sweep for all admitted requests.
 * *active receives the number of requests still in flight. */
int bcx_rq_step(bcx_rq *rq, int32_t *active);
/* 1 and result copied (bc_out[q]) if done, 0 if still running. */
int bcx_rq_poll(bcx_rq *rq, int64_t id, double *bc_out);
int bcx_rq_bytes_needed(const bcx_rq *rq, int32_t q, size_t *bytes);
/* Device memory the engine holds now (fixed part plus admitted requests) and
 * its budget. */
int bcx_rq_memory(const bcx_rq *rq, size_t *used, size_t *budget);

#ifdef __cplusplus
}
#endif
#endif /* BCX_H_ */
