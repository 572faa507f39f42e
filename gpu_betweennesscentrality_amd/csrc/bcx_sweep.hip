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

// Lane state is one record per (tile, node): 64 distances, then 64 path
// counts, then 64 dependencies (1280 bytes), so the rows a node reads of an
// ancestor or descendant are one contiguous region of memory.
constexpr int64_t kRec = kWave * (4 + 8 + 8);
__device__ __forceinline__ int64_t lix_(int nn, int t, int v, int /*lane*/) {
  return ((int64_t)t * nn + v) * kRec;
}
__device__ __forceinline__ int &SD(const SweepWork &sw, int64_t r, int lane) {
  return ((int *)(sw.state + r))[lane];
}
__device__ __forceinline__ uint64_t &SS(const SweepWork &sw, int64_t r, int lane) {
  return ((uint64_t *)(sw.state + r + kWave * 4))[lane];
}
__device__ __forceinline__ double &SDD(const SweepWork &sw, int64_t r, int lane) {
  return ((double *)(sw.state + r + kWave * 12))[lane];
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
// Entries are taken four at a time so their row loads are in flight together.
__device__ __forceinline__ void fwd_bag(const DevTree &tr, const SweepWork &sw, const Lane &L,
                                        int e, int e_end, int stride, int &best, uint64_t &c) {
  const int n = tr.n;
  for (; e < e_end; e += 4 * stride) {
    int cand[4];
    uint64_t m[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int eq = e + q * stride;
      const bool ok = eq < e_end;
      const int ec = ok ? eq : e;
      const int64_t a = lix_(tr.n, L.t, tr.bag_v[ec], L.lane);
      cand[q] = ok ? SD(sw, a, L.lane) + tr.bag_len[ec] : 0x7fffffff;
      m[q] = ok ? SS(sw, a, L.lane) : 0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int eq = e + q * stride;
      if (eq >= e_end) break;
      if (cand[q] < best) {
        best = cand[q];
        c = mul_sat(m[q], tr.bag_cnt[eq]);
      } else if (cand[q] == best) {
        c = add_sat(c, mul_sat(m[q], tr.bag_cnt[eq]));
      }
    }
  }
}

// Backward pulls of node v over inverse-bag entries [k, k_end) with a stride,
// four at a time; writes the tile's flow on each entry and returns the lane's
// dependency part.
__device__ __forceinline__ double bwd_entries(const DevTree &tr, const SweepWork &sw,
                                              const Lane &L, int k, int k_end, int stride,
                                              int dv, double sv) {
  const int n = tr.n;
  double acc = 0.0;
  for (; k < k_end; k += 4 * stride) {
    int dx[4], ee[4], len[4];
    uint64_t sx[4];
    double ddx[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kq = k + q * stride;
      const int kc = kq < k_end ? kq : k;
      const int x = tr.ibag_x[kc];
      ee[q] = tr.ibag_e[kc];
      len[q] = tr.bag_len[ee[q]];
      const int64_t ax = lix_(tr.n, L.t, x, L.lane);
      dx[q] = SD(sw, ax, L.lane);
      sx[q] = SS(sw, ax, L.lane);
      ddx[q] = SDD(sw, ax, L.lane);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kq = k + q * stride;
      if (kq >= k_end) break;
      const double f = dv + len[q] == dx[q]
                           ? (double)tr.bag_cnt[ee[q]] * sv * (1.0 + ddx[q]) / (double)sx[q]
                           : 0.0;
      acc += f;
      const double fs = wave_sum_dpp(L.valid ? f : 0.0);
      if (L.lane == 0) sw.phi_t[(int64_t)L.t * tr.bags + ee[q]] = fs;
    }
  }
  return acc;
}


__device__ __forceinline__ int rl(int x, int j) { return __builtin_amdgcn_readlane(x, j); }
__device__ __forceinline__ uint64_t rl64(uint64_t x, int j) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), j);
  return ((uint64_t)hi << 32) | lo;
}

constexpr int kStage = 4;  // bag entries per node staged in registers

constexpr int kNodes = 2;  // nodes of a level processed together by one wavefront

// Forward over the nodes of one level assigned to this wavefront (nodes
// wave, wave + 16, ...).  Lane j first loads node j's metadata and its first
// kStage bag entries; the wavefront then handles kNodes nodes per step with
// all their row loads in flight together.
__device__ __forceinline__ void fwd_level_wave(const DevIndex &ix, const DevTree &tr,
                                               const SweepWork &sw, const Lane &L, int wave,
                                               int lo, int cnt, int &ovf) {
  const int n = tr.n;
  const int mine = (cnt - wave + kSweepWaves - 1) / kSweepWaves;
  for (int base = 0; base < mine; base += kWave) {
    const int jn = min(kWave, mine - base);
    int sv = 0, se0 = 0, se1 = 0, ssub = 0, sdep = 0;
    int su[kStage], slen[kStage];
    uint64_t sc[kStage];
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
      su[r] = 0;
      slen[r] = 0;
      sc[r] = 0;
    }
    if (L.lane < jn) {
      sv = tr.lvl_v[lo + wave + kSweepWaves * (base + L.lane)];
      se0 = tr.bag_off[sv];
      se1 = tr.bag_off[sv + 1];
      ssub = tr.subtree[sv];
      sdep = ix.depth[sv];
#pragma unroll
      for (int r = 0; r < kStage; ++r)
        if (se0 + r < se1) {
          su[r] = tr.bag_v[se0 + r];
          slen[r] = tr.bag_len[se0 + r];
          sc[r] = tr.bag_cnt[se0 + r];
        }
    }
    for (int j0 = 0; j0 < jn; j0 += kNodes) {
      int cand[kNodes][kStage];
      uint64_t m[kNodes][kStage];
#pragma unroll
      for (int u = 0; u < kNodes; ++u) {
        const int j = min(j0 + u, jn - 1);
        const int nb = min(kStage, rl(se1, j) - rl(se0, j));
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
          const int64_t a = lix_(n, L.t, rl(su[r], j), L.lane);
          cand[u][r] = r < nb ? SD(sw, a, L.lane) + rl(slen[r], j) : 0x7fffffff;
          m[u][r] = r < nb ? SS(sw, a, L.lane) : 0;
        }
      }
#pragma unroll
      for (int u = 0; u < kNodes; ++u) {
        const int j = j0 + u;
        if (j >= jn) break;
        const int v = rl(sv, j), e0 = rl(se0, j), e1 = rl(se1, j);
        const int nb = min(kStage, e1 - e0);
        int best = 0x7fffffff;
        uint64_t c = 0;
        if (v <= L.s && L.s < v + rl(ssub, j)) {  // v is an ancestor of s: up-label
          const int64_t o = L.rs + rl(sdep, j);
          best = ix.dis[o];
          c = ix.cnt[o];
        }
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
          if (r >= nb) break;
          const uint64_t cr = rl64(sc[r], j);
          if (cand[u][r] < best) {
            best = cand[u][r];
            c = mul_sat(m[u][r], cr);
          } else if (cand[u][r] == best) {
            c = add_sat(c, mul_sat(m[u][r], cr));
          }
        }
        if (e1 - e0 > kStage) fwd_bag(tr, sw, L, e0 + kStage, e1, 1, best, c);
        const int64_t a = lix_(n, L.t, v, L.lane);
        SD(sw, a, L.lane) = best;
        SS(sw, a, L.lane) = c;
        ovf |= L.valid && c == ~0ull;
      }
    }
  }
}

// Backward over the nodes of one level assigned to this wavefront, with the
// same staging and kNodes nodes per step.
__device__ __forceinline__ void bwd_level_wave(const DevTree &tr, const SweepWork &sw,
                                               const Lane &L, int wave, int lo, int cnt) {
  const int n = tr.n;
  const int mine = (cnt - wave + kSweepWaves - 1) / kSweepWaves;
  for (int base = 0; base < mine; base += kWave) {
    const int jn = min(kWave, mine - base);
    int sv = 0, sk0 = 0, sk1 = 0;
    int sx[kStage], se[kStage], slen[kStage];
    uint64_t sc[kStage];
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
      sx[r] = 0;
      se[r] = 0;
      slen[r] = 0;
      sc[r] = 0;
    }
    if (L.lane < jn) {
      sv = tr.lvl_v[lo + wave + kSweepWaves * (base + L.lane)];
      sk0 = tr.ibag_off[sv];
      sk1 = tr.ibag_off[sv + 1];
#pragma unroll
      for (int r = 0; r < kStage; ++r)
        if (sk0 + r < sk1) {
          sx[r] = tr.ibag_x[sk0 + r];
          se[r] = tr.ibag_e[sk0 + r];
          slen[r] = tr.bag_len[se[r]];
          sc[r] = tr.bag_cnt[se[r]];
        }
    }
    for (int j0 = 0; j0 < jn; j0 += kNodes) {
      int dv[kNodes];
      uint64_t svu[kNodes];
      int dx[kNodes][kStage];
      uint64_t sxx[kNodes][kStage];
      double ddx[kNodes][kStage];
#pragma unroll
      for (int u = 0; u < kNodes; ++u) {
        const int j = min(j0 + u, jn - 1);
        const int nb = min(kStage, rl(sk1, j) - rl(sk0, j));
        const int64_t av = lix_(n, L.t, rl(sv, j), L.lane);
        dv[u] = SD(sw, av, L.lane);
        svu[u] = SS(sw, av, L.lane);
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
          const int64_t ax = lix_(n, L.t, rl(sx[r], j), L.lane);
          dx[u][r] = r < nb ? SD(sw, ax, L.lane) : -1;
          sxx[u][r] = r < nb ? SS(sw, ax, L.lane) : 1;
          ddx[u][r] = r < nb ? SDD(sw, ax, L.lane) : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < kNodes; ++u) {
        const int j = j0 + u;
        if (j >= jn) break;
        const int v = rl(sv, j), k0 = rl(sk0, j), k1 = rl(sk1, j);
        const int nb = min(kStage, k1 - k0);
        const double svv = (double)svu[u];
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
          if (r >= nb) break;
          const int e = rl(se[r], j);
          const double f = dv[u] + rl(slen[r], j) == dx[u][r]
                               ? (double)rl64(sc[r], j) * svv * (1.0 + ddx[u][r]) /
                                     (double)sxx[u][r]
                               : 0.0;
          acc += f;
          const double fs = wave_sum_dpp(L.valid ? f : 0.0);
          if (L.lane == 0) sw.phi_t[(int64_t)L.t * tr.bags + e] = fs;
        }
        if (k1 - k0 > kStage) acc += bwd_entries(tr, sw, L, k0 + kStage, k1, 1, dv[u], svv);
        const int64_t av = lix_(n, L.t, v, L.lane);
        SDD(sw, av, L.lane) = acc;
        const double ns = wave_sum_dpp((L.valid && L.s != v) ? acc : 0.0);
        if (L.lane == 0) sw.node_t[(int64_t)L.t * n + v] = ns;
      }
    }
  }
}

// One workgroup per tile of 64 sources runs the whole sweep for its tile:
// forward over depth levels, backward over depth levels, then the ascending
// states.  Levels narrower than the workgroup's 16 wavefronts are processed
// node by node with the node's entries split over all wavefronts and the
// partial results merged in wavefront order, so results are deterministic.
__global__ __launch_bounds__(kSweepThreads) void sweep_tile_kernel(DevIndex ix, DevTree tr,
                                                                   SweepWork sw, int nsrc,
                                                                   int *ovf_flag,
                                                                   uint64_t *stamps) {
  __shared__ double sh_acc[kSweepWaves][kWave];
  __shared__ int sh_best[kSweepWaves][kWave];
  __shared__ uint64_t sh_cnt[kSweepWaves][kWave];
  const Lane L = lane_setup(ix, sw, blockIdx.x, nsrc);
  const int wave = threadIdx.x / kWave;
  const int n = tr.n, H = tr.height;
  int ovf = 0;
  // Optional phase timestamps (constant 100 MHz clock), one row per tile.
  if (stamps && threadIdx.x == 0) stamps[4 * blockIdx.x] = __builtin_amdgcn_s_memrealtime();

  // ---- forward: distance and path count ----
  for (int k = 0; k < H; ++k) {
    const int lo = tr.lvl_off[k], cnt = tr.lvl_off[k + 1] - lo;
    if (cnt >= kSweepWaves) {
      fwd_level_wave(ix, tr, sw, L, wave, lo, cnt, ovf);
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
          const int64_t a = lix_(tr.n, L.t, v, L.lane);
          SD(sw, a, L.lane) = b;
          SS(sw, a, L.lane) = cc;
          ovf |= L.valid && cc == ~0ull;
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }

  if (stamps && threadIdx.x == 0) stamps[4 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();

  // ---- backward: dependencies, shortcut flows, node flows ----
  for (int k = H - 1; k >= 0; --k) {
    const int lo = tr.lvl_off[k], cnt = tr.lvl_off[k + 1] - lo;
    if (cnt >= kSweepWaves) {
      bwd_level_wave(tr, sw, L, wave, lo, cnt);
    } else {
      for (int q = 0; q < cnt; ++q) {
        const int v = __builtin_amdgcn_readfirstlane(tr.lvl_v[lo + q]);
        const int64_t av = lix_(tr.n, L.t, v, L.lane);
        sh_acc[wave][L.lane] = bwd_entries(tr, sw, L, tr.ibag_off[v] + wave, tr.ibag_off[v + 1],
                                           kSweepWaves, SD(sw, av, L.lane), (double)SS(sw, av, L.lane));
        __syncthreads();
        if (wave == 0) {
          double acc = 0.0;
          for (int w = 0; w < kSweepWaves; ++w) acc += sh_acc[w][L.lane];
          SDD(sw, av, L.lane) = acc;
          const double ns = wave_sum_dpp((L.valid && L.s != v) ? acc : 0.0);
          if (L.lane == 0) sw.node_t[(int64_t)L.t * tr.n + v] = ns;
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }

  if (stamps && threadIdx.x == 0) stamps[4 * blockIdx.x + 2] = __builtin_amdgcn_s_memrealtime();

  // ---- ascending states: the union of the tile's root paths in top-down
  // order, each node's bag entries split over all wavefronts and merged in
  // wavefront order.  Once node x is done, its slots hold the ascending state:
  // S(x) := sigma_A(x) (the up-label count where the up-label distance is
  // tight, else 0) and dD(x) := dA(x), with
  //   dA(x) = sigma_A(x) * (1 + dD(x)) / sigma(x)
  //         + sum over tight bag entries (x -> y) c * sigma_A(x) * dA(y) / sigma_A(y).
  // Later nodes read only their ancestors' ascending state, and a lane whose
  // source is not below x never reads x again.
  {
    const int c0 = sw.cand_off[L.t], c1 = sw.cand_off[L.t + 1];
    for (int kk = c0; kk < c1; ++kk) {
      const int x = __builtin_amdgcn_readfirstlane(sw.cand[kk]);
      const bool anc = x <= L.s && L.s < x + tr.subtree[x];
      const int64_t ax = lix_(n, L.t, x, L.lane);
      const int dx = SD(sw, ax, L.lane);
      uint64_t sa = 0;
      if (anc) {
        const int64_t o = L.rs + ix.depth[x];
        if (ix.dis[o] == dx) sa = ix.cnt[o];
      }
      const double sax = (double)sa;
      double up = 0.0;
      if (__any(sa != 0)) {
        const int e1 = tr.bag_off[x + 1];
        for (int e = tr.bag_off[x] + wave; e < e1; e += 4 * kSweepWaves) {
          double f[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int eq = e + q * kSweepWaves;
            const int ec = eq < e1 ? eq : e;
            const int64_t ay = lix_(n, L.t, tr.bag_v[ec], L.lane);
            const int dy = SD(sw, ay, L.lane);
            const uint64_t say = SS(sw, ay, L.lane);  // sigma_A(y), y already processed
            const double day = SDD(sw, ay, L.lane);   // dA(y)
            f[q] = (eq < e1 && sa != 0 && say != 0 && dx + tr.bag_len[ec] == dy)
                       ? (double)tr.bag_cnt[ec] * sax * day / (double)say
                       : 0.0;
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int eq = e + q * kSweepWaves;
            if (eq >= e1) break;
            up += f[q];
            const double fs = wave_sum_dpp(L.valid ? f[q] : 0.0);
            if (L.lane == 0) sw.phi_t[(int64_t)L.t * tr.bags + eq] += fs;
          }
        }
      }
      sh_acc[wave][L.lane] = up;
      __syncthreads();
      if (wave == 0) {
        double u = 0.0;
        for (int w = 0; w < kSweepWaves; ++w) u += sh_acc[w][L.lane];
        const double ddx = SDD(sw, ax, L.lane);
        const double sx = (double)SS(sw, ax, L.lane);
        SDD(sw, ax, L.lane) = sa != 0 ? sax * (1.0 + ddx) / sx + u : 0.0;
        SS(sw, ax, L.lane) = sa;
        const double ns = wave_sum_dpp((L.valid && x != L.s) ? u : 0.0);
        if (L.lane == 0) sw.node_t[(int64_t)L.t * n + x] += ns;
      }
      __syncthreads();
    }
  }
  if (stamps && threadIdx.x == 0) stamps[4 * blockIdx.x + 3] = __builtin_amdgcn_s_memrealtime();
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
  void *bufs[] = {sw.state, sw.node_t, sw.phi_t,
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
  HIP_OK(hipMalloc(&sw.state, T * n * W * 20));
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
  sw.tstride = tiles;
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
    uint64_t *stamps = nullptr;
    if (timed) HIP_OK(hipMalloc(&stamps, (size_t)bt * 4 * 8));
    sweep_tile_kernel<<<bt, kSweepThreads, 0, c->stream>>>(c->ix, c->tree, sw, nsrc, c->d_ovf,
                                                           stamps);
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
      std::vector<uint64_t> st((size_t)bt * 4);
      HIP_OK(hipMemcpy(st.data(), stamps, st.size() * 8, hipMemcpyDeviceToHost));
      HIP_OK(hipFree(stamps));
      double f = 0, b2 = 0, as = 0;
      for (int t = 0; t < bt; ++t) {
        f += (double)(st[4 * t + 1] - st[4 * t]);
        b2 += (double)(st[4 * t + 2] - st[4 * t + 1]);
        as += (double)(st[4 * t + 3] - st[4 * t + 2]);
      }
      c->timing.sweep_fwd_ms += f / bt * 1e-5;  // 100 MHz ticks
      c->timing.sweep_bwd_ms += b2 / bt * 1e-5;
      c->timing.sweep_asc_ms += as / bt * 1e-5;
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
