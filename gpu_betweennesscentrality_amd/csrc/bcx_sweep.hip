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

__device__ __forceinline__ int64_t lix(int t, int n, int v, int lane) {
  return ((int64_t)t * n + v) * kWave + lane;
}

struct Lane {
  int t, lane, s;
  bool valid;
  int64_t rs;
};

__device__ __forceinline__ Lane lane_setup(const DevIndex &ix, const SweepWork &sw, int t,
                                           int nsrc) {
  Lane L;
  L.t = t;
  L.lane = threadIdx.x & (kWave - 1);
  const int slot = t * kWave + L.lane;
  L.valid = slot < nsrc;
  L.s = sw.tile_src[slot];  // padded with a valid source beyond nsrc
  L.rs = ix.row_off[L.s];
  return L;
}

// Sum of a double over the 64 lanes of a wavefront with DPP lane moves
// (quad swaps, row rotations, row broadcasts), in one fixed order; the total
// is taken from lane 63.  All 64 lanes must be active.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, kCtrl, kRowMask, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, kCtrl, kRowMask, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_move<0xb1>(v);        // quad_perm [1,0,3,2]
  v += dpp_move<0x4e>(v);        // quad_perm [2,3,0,1]: quad sums
  v += dpp_move<0x124>(v);       // row_ror:4
  v += dpp_move<0x128>(v);       // row_ror:8: row sums in every lane
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

constexpr int kSweepThreads = 1024;
constexpr int kSweepWaves = kSweepThreads / kWave;

// Forward candidates of node v from bag entries [e, e_end) with a stride.
__device__ __forceinline__ void fwd_bag(const DevTree &tr, const SweepWork &sw, const Lane &L,
                                        int e, int e_end, int stride, int &best, uint64_t &c) {
  const int n = tr.n;
  for (; e < e_end; e += stride) {
    const int64_t a = lix(L.t, n, tr.bag_v[e], L.lane);
    const int cand = sw.D[a] + tr.bag_len[e];
    const uint64_t su = sw.S[a];
    if (cand < best) {
      best = cand;
      c = mul_sat(su, tr.bag_cnt[e]);
    } else if (cand == best) {
      c = add_sat(c, mul_sat(su, tr.bag_cnt[e]));
    }
  }
}

// Backward pulls of node v over inverse-bag entries [k, k_end) with a stride;
// writes the tile's flow on each entry and returns the lane's dependency part.
__device__ __forceinline__ double bwd_entries(const DevTree &tr, const SweepWork &sw,
                                              const Lane &L, int k, int k_end, int stride,
                                              int dv, double sv) {
  const int n = tr.n;
  double acc = 0.0;
  for (; k < k_end; k += stride) {
    const int x = tr.ibag_x[k], e = tr.ibag_e[k];
    const int64_t ax = lix(L.t, n, x, L.lane);
    const int dx = sw.D[ax];
    const uint64_t sx = sw.S[ax];
    const double ddx = sw.dD[ax];
    const double f =
        dv + tr.bag_len[e] == dx ? (double)tr.bag_cnt[e] * sv * (1.0 + ddx) / (double)sx : 0.0;
    acc += f;
    const double fs = wave_sum_dpp(L.valid ? f : 0.0);
    if (L.lane == 0) sw.phi_t[(int64_t)L.t * tr.bags + e] = fs;
  }
  return acc;
}

// One workgroup per tile of 64 sources runs the whole sweep for its tile:
// forward over depth levels, backward over depth levels, then the ascending
// states.  Levels narrower than the workgroup's 16 wavefronts are processed
// node by node with the node's entries split over all wavefronts and the
// partial results merged in wavefront order, so results are deterministic.
__global__ __launch_bounds__(kSweepThreads) void sweep_tile_kernel(DevIndex ix, DevTree tr,
                                                                   SweepWork sw, int nsrc,
                                                                   int *ovf_flag) {
  __shared__ double sh_acc[kSweepWaves][kWave];
  __shared__ int sh_best[kSweepWaves][kWave];
  __shared__ uint64_t sh_cnt[kSweepWaves][kWave];
  const Lane L = lane_setup(ix, sw, blockIdx.x, nsrc);
  const int wave = threadIdx.x / kWave;
  const int n = tr.n, H = tr.height;
  int ovf = 0;

  // ---- forward: distance and path count ----
  for (int k = 0; k < H; ++k) {
    const int lo = tr.lvl_off[k], cnt = tr.lvl_off[k + 1] - lo;
    if (cnt >= kSweepWaves) {
      for (int q = wave; q < cnt; q += kSweepWaves) {
        const int v = __builtin_amdgcn_readfirstlane(tr.lvl_v[lo + q]);
        int best = 0x7fffffff;
        uint64_t c = 0;
        if (v <= L.s && L.s < v + tr.subtree[v]) {  // v is an ancestor of s: up-label
          const int64_t o = L.rs + ix.depth[v];
          best = ix.dis[o];
          c = ix.cnt[o];
        }
        fwd_bag(tr, sw, L, tr.bag_off[v], tr.bag_off[v + 1], 1, best, c);
        const int64_t a = lix(L.t, n, v, L.lane);
        sw.D[a] = best;
        sw.S[a] = c;
        ovf |= L.valid && c == ~0ull;
      }
    } else {
      for (int q = 0; q < cnt; ++q) {
        const int v = __builtin_amdgcn_readfirstlane(tr.lvl_v[lo + q]);
        int best = 0x7fffffff;
        uint64_t c = 0;
        if (wave == 0 && v <= L.s && L.s < v + tr.subtree[v]) {
          const int64_t o = L.rs + ix.depth[v];
          best = ix.dis[o];
          c = ix.cnt[o];
        }
        fwd_bag(tr, sw, L, tr.bag_off[v] + wave, tr.bag_off[v + 1], kSweepWaves, best, c);
        sh_best[wave][L.lane] = best;
        sh_cnt[wave][L.lane] = c;
        __syncthreads();
        if (wave == 0) {
          int b = sh_best[0][L.lane];
          uint64_t cc = sh_cnt[0][L.lane];
          for (int w = 1; w < kSweepWaves; ++w) {
            const int bw = sh_best[w][L.lane];
            if (bw < b) {
              b = bw;
              cc = sh_cnt[w][L.lane];
            } else if (bw == b) {
              cc = add_sat(cc, sh_cnt[w][L.lane]);
            }
          }
          const int64_t a = lix(L.t, n, v, L.lane);
          sw.D[a] = b;
          sw.S[a] = cc;
          ovf |= L.valid && cc == ~0ull;
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }

  // ---- backward: dependencies, shortcut flows, node flows ----
  for (int k = H - 1; k >= 0; --k) {
    const int lo = tr.lvl_off[k], cnt = tr.lvl_off[k + 1] - lo;
    if (cnt >= kSweepWaves) {
      for (int q = wave; q < cnt; q += kSweepWaves) {
        const int v = __builtin_amdgcn_readfirstlane(tr.lvl_v[lo + q]);
        const int64_t av = lix(L.t, n, v, L.lane);
        const double acc = bwd_entries(tr, sw, L, tr.ibag_off[v], tr.ibag_off[v + 1], 1,
                                       sw.D[av], (double)sw.S[av]);
        sw.dD[av] = acc;
        const double ns = wave_sum_dpp((L.valid && L.s != v) ? acc : 0.0);
        if (L.lane == 0) sw.node_t[(int64_t)L.t * n + v] = ns;
      }
    } else {
      for (int q = 0; q < cnt; ++q) {
        const int v = __builtin_amdgcn_readfirstlane(tr.lvl_v[lo + q]);
        const int64_t av = lix(L.t, n, v, L.lane);
        sh_acc[wave][L.lane] = bwd_entries(tr, sw, L, tr.ibag_off[v] + wave, tr.ibag_off[v + 1],
                                           kSweepWaves, sw.D[av], (double)sw.S[av]);
        __syncthreads();
        if (wave == 0) {
          double acc = 0.0;
          for (int w = 0; w < kSweepWaves; ++w) acc += sh_acc[w][L.lane];
          sw.dD[av] = acc;
          const double ns = wave_sum_dpp((L.valid && L.s != v) ? acc : 0.0);
          if (L.lane == 0) sw.node_t[(int64_t)L.t * n + v] = ns;
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }

  // ---- ascending states (one wavefront; the union of the tile's root paths
  // in top-down order).  dA(x) replaces dD(x) in place once dD(x) is read:
  // later nodes read dA of their ancestors, and a lane whose source is not
  // below x never reads x again.
  //   dA(x) = sigma_A(x) * (1 + dD(x)) / sigma(x)
  //         + sum over tight bag entries (x -> y) c * sigma_A(x) * dA(y) / sigma_A(y),
  // sigma_A = the up-label count where the up-label distance is tight.
  if (wave == 0) {
    const int c0 = sw.cand_off[L.t], c1 = sw.cand_off[L.t + 1];
    for (int kk = c0; kk < c1; ++kk) {
      const int x = __builtin_amdgcn_readfirstlane(sw.cand[kk]);
      const bool anc = x <= L.s && L.s < x + tr.subtree[x];
      const int64_t ax = lix(L.t, n, x, L.lane);
      const int dx = sw.D[ax];
      const double ddx = sw.dD[ax];
      uint64_t sa = 0;
      if (anc) {
        const int64_t o = L.rs + ix.depth[x];
        if (ix.dis[o] == dx) sa = ix.cnt[o];
      }
      const double sax = (double)sa;
      double up = 0.0;
      if (__any(sa != 0)) {
        for (int e = tr.bag_off[x]; e < tr.bag_off[x + 1]; ++e) {
          double f = 0.0;
          if (sa != 0) {
            const int y = tr.bag_v[e];
            const int64_t ay = lix(L.t, n, y, L.lane);
            const int dy = sw.D[ay];
            if (dx + tr.bag_len[e] == dy) {
              const int64_t oy = L.rs + ix.depth[y];
              if (ix.dis[oy] == dy) f = (double)tr.bag_cnt[e] * sax * sw.dD[ay] / (double)ix.cnt[oy];
            }
          }
          up += f;
          const double fs = wave_sum_dpp(L.valid ? f : 0.0);
          if (L.lane == 0) sw.phi_t[(int64_t)L.t * tr.bags + e] += fs;
        }
      }
      sw.dD[ax] = sa != 0 ? sax * (1.0 + ddx) / (double)sw.S[ax] + up : 0.0;
      const double ns = wave_sum_dpp((L.valid && x != L.s) ? up : 0.0);
      if (L.lane == 0) sw.node_t[(int64_t)L.t * n + x] += ns;
    }
  }
  if (ovf) atomicOr(ovf_flag, 1);
}

// node[v] += sum over tiles of node_t, phi[e] += sum over tiles of phi_t,
// tiles in increasing order.
__global__ __launch_bounds__(256) void sweep_tile_reduce_kernel(int n, int64_t B, int tiles,
                                                                SweepWork sw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double acc = sw.node[i];
    for (int t = 0; t < tiles; ++t) acc += sw.node_t[(int64_t)t * n + i];
    sw.node[i] = acc;
  } else if (i < n + B) {
    const int64_t e = i - n;
    double acc = sw.phi[e];
    for (int t = 0; t < tiles; ++t) acc += sw.phi_t[(int64_t)t * B + e];
    sw.phi[e] = acc;
  }
}

// Unpack one depth level, top-down.  For node z and each tight pair (a, b) of
// its bag (a path a-z-b as short as shortcut (a, b)), z carries the share
// c(a,z) c(z,b) / c(a,b) of that shortcut's flow, and the same amount moves
// onto z's shortcuts (a, z) and (z, b) for deeper levels: the
// restricted-path counting of Reduce (TreeDecomp.cpp:157-193), reversed.
// Step 1: the share of every tight pair of the level.
__global__ __launch_bounds__(256) void sweep_pair_flow_kernel(DevTree tr, SweepWork sw,
                                                              int q_lo, int q_cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= q_cnt) return;
  const int q = q_lo + k;
  sw.pf[q] = sw.phi[tr.tp_ab[q]] * tr.tp_frac[q];
}

// Step 2: per node of the level, its interior flow and its entries' new flow,
// summed over pairs in increasing pair order.
__global__ __launch_bounds__(256) void sweep_pair_apply_kernel(DevTree tr, SweepWork sw,
                                                               int lvl_lo, int lvl_cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= lvl_cnt) return;
  const int z = tr.lvl_v[lvl_lo + k];
  double acc = 0.0;
  for (int i = tr.zp_off[z]; i < tr.zp_off[z + 1]; ++i) acc += sw.pf[tr.zp_pair[i]];
  sw.interior[z] = acc;
  for (int e = tr.bag_off[z]; e < tr.bag_off[z + 1]; ++e) {
    double f = 0.0;
    for (int i = tr.ep_off[e]; i < tr.ep_off[e + 1]; ++i) f += sw.pf[tr.ep_pair[i]];
    if (tr.ep_off[e + 1] > tr.ep_off[e]) sw.phi[e] += f;
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
  void *bufs[] = {sw.D,   sw.S,        sw.dD,       sw.node_t, sw.phi_t,
                  sw.node, sw.phi,     sw.interior, sw.tile_src, sw.cand, sw.cand_off,
                  sw.pf};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  sw = SweepWork{};
  sweep_tiles = 0;
}

void bcx_ctx::ensure_sweep(int tiles) {
  if (tiles <= sweep_tiles) return;
  free_sweep();
  const size_t n = (size_t)ix.n, B = (size_t)tree.bags, T = (size_t)tiles, W = kWave;
  HIP_OK(hipMalloc(&sw.D, T * n * W * 4));
  HIP_OK(hipMalloc(&sw.S, T * n * W * 8));
  HIP_OK(hipMalloc(&sw.dD, T * n * W * 8));
  HIP_OK(hipMalloc(&sw.node_t, T * n * 8));
  HIP_OK(hipMalloc(&sw.phi_t, T * std::max<size_t>(B, 1) * 8));
  HIP_OK(hipMalloc(&sw.node, n * 8));
  HIP_OK(hipMalloc(&sw.phi, std::max<size_t>(B, 1) * 8));
  HIP_OK(hipMalloc(&sw.interior, n * 8));
  HIP_OK(hipMalloc(&sw.pf, std::max<size_t>((size_t)tree.pairs, 1) * 8));
  HIP_OK(hipMalloc(&sw.tile_src, T * W * 4));
  HIP_OK(hipMalloc(&sw.cand_off, (T + 1) * 4));
  sw.cand_cap = (int)std::min<size_t>(T * ((size_t)tree.height + W), (size_t)1 << 30);
  HIP_OK(hipMalloc(&sw.cand, (size_t)sw.cand_cap * 4));
  sweep_tiles = tiles;
}

namespace bcx_dev {

void run_sweep(bcx_ctx *c, const int32_t *sources, int32_t count, double *bc_inout,
               uint32_t *overflow) {
  const int n = c->ix.n;
  const int64_t B = c->tree.bags;
  const int H = c->tree.height;
  std::vector<int32_t> pre((size_t)count);
  for (int32_t k = 0; k < count; ++k) {
    if (sources[k] < 0 || sources[k] >= n) throw bcx::Error{BCX_EINVAL, "source out of range"};
    pre[(size_t)k] = c->iperm[(size_t)sources[k]];
  }
  std::sort(pre.begin(), pre.end());
  const int total_tiles = (count + kWave - 1) / kWave;
  int tiles_per_batch = std::max(1, std::min(c->tn.batch / kWave, std::max(1, total_tiles)));
  {
    // Keep the per-tile state within ~70% of the device memory not already
    // held by this context's sweep buffers.
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    const size_t per_tile = (size_t)n * kWave * 20 + (size_t)n * 8 + (size_t)B * 8;
    const size_t held = (size_t)c->sweep_tiles * per_tile;
    const size_t cap = (size_t)((double)(fr + held) * 0.7) / std::max<size_t>(per_tile, 1);
    if (cap < 1) throw bcx::Error{BCX_ENOMEM, "not enough device memory for one sweep tile"};
    tiles_per_batch = (int)std::min<size_t>((size_t)tiles_per_batch, cap);
  }
  c->ensure_sweep(tiles_per_batch);
  SweepWork sw = c->sw;
  std::vector<double> bc_pre((size_t)n);
  for (int32_t v = 0; v < n; ++v) bc_pre[(size_t)c->iperm[(size_t)v]] = bc_inout[v];
  const int zero = 0;
  HIP_OK(hipMemcpyAsync(c->d_ovf, &zero, 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(c->d_bc, bc_pre.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  zero_f64_kernel<<<blocks_for(n, 256), 256, 0, c->stream>>>(sw.node, n);
  if (B > 0) zero_f64_kernel<<<blocks_for(B, 256), 256, 0, c->stream>>>(sw.phi, B);
  const auto &lvl = c->h_lvl_off;
  const auto &par = c->h_parent;
  const auto &sub = c->h_subtree;
  std::vector<int32_t> tsrc, cand, cand_off, path;
  const bool timed = c->profiling;
  for (int tb = 0; tb < total_tiles; tb += tiles_per_batch) {
    const int bt = std::min(tiles_per_batch, total_tiles - tb);
    const int first = tb * kWave;
    const int nsrc = std::min(bt * kWave, count - first);
    // Tile sources (padded with the tile's first source) and the union of
    // their root paths in an order where every ancestor comes first.
    tsrc.assign((size_t)bt * kWave, 0);
    cand.clear();
    cand_off.assign((size_t)bt + 1, 0);
    for (int t = 0; t < bt; ++t) {
      const int t0 = first + t * kWave;
      const int tn = std::min(kWave, count - t0);
      for (int j = 0; j < kWave; ++j) tsrc[(size_t)t * kWave + j] = pre[(size_t)(t0 + (j < tn ? j : 0))];
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
      cand_off[(size_t)t + 1] = (int32_t)cand.size();
    }
    if ((int64_t)cand.size() > c->sw.cand_cap) {
      HIP_OK(hipStreamSynchronize(c->stream));
      HIP_OK(hipFree(c->sw.cand));
      c->sw.cand_cap = (int)cand.size();
      HIP_OK(hipMalloc(&c->sw.cand, cand.size() * 4));
      sw.cand = c->sw.cand;
    }
    HIP_OK(hipMemcpyAsync(sw.tile_src, tsrc.data(), tsrc.size() * 4, hipMemcpyHostToDevice,
                          c->stream));
    HIP_OK(hipMemcpyAsync(sw.cand_off, cand_off.data(), cand_off.size() * 4,
                          hipMemcpyHostToDevice, c->stream));
    if (!cand.empty())
      HIP_OK(hipMemcpyAsync(sw.cand, cand.data(), cand.size() * 4, hipMemcpyHostToDevice,
                            c->stream));
    if (timed) HIP_OK(hipEventRecord(c->ev[0], c->stream));
    sweep_tile_kernel<<<bt, kSweepThreads, 0, c->stream>>>(c->ix, c->tree, sw, nsrc, c->d_ovf);
    if (timed) HIP_OK(hipEventRecord(c->ev[1], c->stream));
    if (timed) HIP_OK(hipEventRecord(c->ev[2], c->stream));
    sweep_tile_reduce_kernel<<<blocks_for(n + B, 256), 256, 0, c->stream>>>(n, B, bt, sw);
    HIP_OK(hipGetLastError());
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
    }
    c->timing.scan_launches += 1;
    c->timing.dep_launches += 1;
    // The host vectors are reused by the next batch: wait for the copies.
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  const auto &lpo = c->h_lvl_pair_off;
  for (int k = 0; k < H; ++k) {
    const int lo = lvl[(size_t)k], cnt = lvl[(size_t)k + 1] - lo;
    const int qlo = lpo[(size_t)k], qcnt = lpo[(size_t)k + 1] - qlo;
    if (qcnt == 0) {
      sweep_pair_apply_kernel<<<blocks_for(cnt, 256), 256, 0, c->stream>>>(c->tree, sw, lo, cnt);
      continue;
    }
    sweep_pair_flow_kernel<<<blocks_for(qcnt, 256), 256, 0, c->stream>>>(c->tree, sw, qlo, qcnt);
    sweep_pair_apply_kernel<<<blocks_for(cnt, 256), 256, 0, c->stream>>>(c->tree, sw, lo, cnt);
  }
  sweep_combine_kernel<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, sw, c->d_bc);
  HIP_OK(hipGetLastError());
  c->timing.sources += count;
  int ovf = 0;
  HIP_OK(hipMemcpyAsync(bc_pre.data(), c->d_bc, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(&ovf, c->d_ovf, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  for (int32_t v = 0; v < n; ++v) bc_inout[v] = bc_pre[(size_t)c->iperm[(size_t)v]];
  if (overflow) *overflow = (uint32_t)ovf;
}

}  // namespace bcx_dev
