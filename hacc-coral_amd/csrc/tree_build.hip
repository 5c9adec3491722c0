// tree_build.hip -- level-synchronous, bit-exact RCB tree build for gfx950.
//
// Restates RCBForceTree<1>::createRCBForceTree and its helpers
// (RCBForceTree.cxx:519-569, 623-918; BGQCM.c:181-212) as breadth-first
// passes over SoA float32 arrays in HBM:
//
//   per level:  cm (tight bbox + exact double centre of mass)  -> classify
//               (leaf / split dim / pivot)  -> partition of every splitting
//               node (flag count, in-node scan, rank, swap-chain resolve,
//               scatter to the other buffer)  -> children of the next level
//   after:      pseudo-particles bottom-up, canonical (unthreaded allocation
//               order) numbering top-down, one payload gather.
//
// Compiled with -ffp-contract=off: every float operation that shapes the
// tree (pptdr, tdr, split-length compares) rounds exactly as the reference's
// x86 build does.
#include "hsr_internal.h"
#include "softdouble.h"
#include "wave.h"

#include <algorithm>
#include <cstdio>

namespace {

constexpr int CM_BLK = 1024;        // elements per block of a big node
constexpr int CM_PER = CM_BLK / 64; // consecutive addends each lane composes
constexpr int64_t CM_SMALL = 2048;  // nodes up to this size: one wave each
constexpr int PB = 1024;            // partition block (256 threads x 4)
constexpr int MAXHOP = 256;         // swap-chain hops before pointer jumping (default)

struct Counters {
  int64_t cm_blocks;      // big-node cm blocks allocated this level
  int64_t part_blocks;    // partition blocks allocated this level
  int64_t nsplit;         // non-degenerate splits this level
  int64_t active_next;    // particles in the next level's nodes
  int64_t unresolved;     // swap chains longer than MAXHOP
  int64_t cm_fail;        // small-node sums left to the ordered path
  int64_t pad[2];
  int64_t next_big;       // children of this level with count > CM_SMALL
  int64_t next_small;     // and with count <= CM_SMALL
  int64_t next_leaf;      // and with count <= nd (leaves of the next level)
  int64_t ndegen;         // degenerate splits of this level
};

// ------------------------------------------------------------ helpers
__device__ __forceinline__ float pptdr_dev(const float *xmin, const float *xmax,
                                           const float *xc) {
  float a = xmax[0] - xc[0], b = xmax[1] - xc[1], c = xmax[2] - xc[2];
  float d = xc[0] - xmin[0], e = xc[1] - xmin[1], f = xc[2] - xmin[2];
  // std::min nesting of RCBForceTree.cxx:521-522; std::min(p,q) = q<p ? q : p
  float r = (f < e) ? f : e;
  r = (r < d) ? r : d;
  r = (r < c) ? r : c;
  r = (r < b) ? r : b;
  r = (r < a) ? r : a;
  return r;
}

__device__ __forceinline__ const float *coord(int d, const float *X,
                                              const float *Y, const float *Z) {
  return d == 0 ? X : (d == 1 ? Y : Z);
}

// A finite float as odd * 2^lsb (odd = 0 and lsb = 2^20 for zero), without
// branches: the mantissa with its hidden bit, its trailing zeros shifted out.
__device__ __forceinline__ void f32_odd_lsb(float a, uint32_t &odd, int32_t &lsb) {
  const uint32_t u = __float_as_uint(a);
  const uint32_t e = (u >> 23) & 255u, fr = u & 0x7fffffu;
  const uint32_t fm = e ? (fr | 0x800000u) : fr;
  const int32_t fe = e ? (int32_t)e - 150 : -149;
  const int tz = __builtin_ctz(fm | 0x80000000u);  // fm < 2^24: bit 31 stops a zero
  odd = fm >> tz;
  lsb = fm ? fe + tz : (1 << 20);
}

// A lane's exact integer sum of nonnegative floats in units of the finest
// lsb t seen so far (one pass: a finer addend rescales the sum).  bad: an
// addend is negative, or the sum reaches 2^53 units (in any finer unit it
// stays there), so no exact integer sum below 2^53 exists.
struct LaneSum {
  uint64_t u = 0;
  int32_t t = 1 << 20;
  bool bad = false;
};

__device__ __forceinline__ void lane_sum_add(LaneSum &S, float a) {
  uint32_t odd;
  int32_t la;
  f32_odd_lsb(a, odd, la);
  const bool nz = odd != 0u;
  const bool finer = nz && la < S.t;
  const uint32_t d = finer ? (uint32_t)(S.t - la) : 0u;
  const bool ovr = finer && S.u != 0 && (d >= 53u || (S.u >> (53u - d)) != 0);
  S.u = finer ? (d >= 53u ? 0ull : S.u << d) : S.u;
  S.t = finer ? la : S.t;
  const uint32_t sh = nz ? (uint32_t)(la - S.t) : 0u;  // la >= t now
  const bool ova = nz && (sh >= 53u || ((uint64_t)odd >> (53u - sh)) != 0);
  S.u += (nz && !ova) ? ((uint64_t)odd << sh) : 0ull;
  S.bad = S.bad || ovr || ova || S.u >= SD_M53 || (nz && a < 0.f);
}

// The wave's sum from its lanes': the finest lsb t over the wave and the
// total in units of 2^t, or bad (no exact integer sum below 2^53 units).
__device__ __forceinline__ uint64_t wave_lane_sum(const LaneSum &S, int32_t &t, bool &bad) {
  t = wave_min_i32(S.t);
  const uint32_t d = (uint32_t)(S.t - t);
  const bool ov = S.u != 0 && (d >= 53u || (S.u >> (53u - d)) != 0);
  const uint64_t u = ov ? 0ull : (d >= 53u ? 0ull : S.u << d);
  const uint64_t tot = wave_sum_u64(u);  // 64 values below 2^53
  bad = __any(S.bad || ov) || tot >= SD_M53;
  return tot;
}

// A block's element range from its node: blocks of bs elements from the
// node's first element, blk0 the node's first block in the level's table.
__device__ __forceinline__ void blk_range(const BNode &n, int64_t b, int64_t bs, int64_t &start,
                                          int &len) {
  const int64_t rel = (b - n.blk0) * bs;
  start = n.offset + rel;
  len = (int)min<int64_t>(bs, n.count - rel);
}

// ------------------------------------------------------------ cm (small)
// With unit masses (every mass exactly 1.0f) the addend w * x is x exactly,
// the mass sum is the count, and M is not read (the partition then stops
// moving masses, so M may hold stale values).
__device__ __forceinline__ float addend(int q, bool unit, const float *X, const float *Y,
                                        const float *Z, const float *M,
                                        int64_t e) {
  if (q == 3) return unit ? 1.f : M[e];
  const float c = q == 0 ? X[e] : (q == 1 ? Y[e] : Z[e]);
  return unit ? c : M[e] * c;
}

// buf[k] = addend q of element off + k for k < n, U loads in flight per lane
// (the staging is latency-bound otherwise).
template <int U>
__device__ __forceinline__ void stage_addends(float *buf, int q, bool unit, const float *X, const float *Y,
                                              const float *Z, const float *M, int64_t off,
                                              int64_t n) {
  const int l = lane_id();
  for (int64_t k0 = 0; k0 < n; k0 += 64 * U) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + u * 64 + l;
      v[u] = k < n ? addend(q, unit, X, Y, Z, M, off + k) : 0.f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + u * 64 + l;
      if (k < n) buf[k] = v[u];
    }
  }
}

// One wave per (node, sum) for nodes with count <= CM_SMALL
// (BGQCM.c:181-212), sum q (x, y, z, mass) = blockIdx.y; the q = 0 wave also
// takes the bbox by wave min/max.  The addends stay in registers and are
// summed exactly when every partial sum is a multiple of the finest lsb
// below 2^53 of them (no rounding anywhere, so the sequential double sum is
// that integer sum; the first test of sd_block).  Sums that round go to
// cm_list for k_cm_small_ordered.  k_cm_finish divides.
constexpr int SM_PER = CM_SMALL / 64;
__global__ void __launch_bounds__(64)
k_cm_small(BNode *nodes, int64_t lb, int64_t le, bool unit, const float *X,
           const float *Y, const float *Z, const float *M, sd_t *cms,
           int32_t *cm_list, Counters *ctr) {
  const int64_t t = lb + blockIdx.x;
  if (t >= le) return;
  const int q = blockIdx.y;
  const int64_t cnt = nodes[t].count;
  if (cnt > CM_SMALL || cnt == 0) return;
  const int64_t off = nodes[t].offset;
  const int l = lane_id();
  if (q == 0) {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t k0 = 0; k0 < cnt; k0 += 64 * 4) {
      float px[4], py[4], pz[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t k = k0 + u * 64 + l;
        const bool v = k < cnt;
        px[u] = v ? X[off + k] : X[off];
        py[u] = v ? Y[off + k] : Y[off];
        pz[u] = v ? Z[off + k] : Z[off];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        mn[0] = fminf(mn[0], px[u]); mx[0] = fmaxf(mx[0], px[u]);
        mn[1] = fminf(mn[1], py[u]); mx[1] = fmaxf(mx[1], py[u]);
        mn[2] = fminf(mn[2], pz[u]); mx[2] = fmaxf(mx[2], pz[u]);
      }
    }
    for (int d = 0; d < 3; ++d) {
      mn[d] = wave_min_f32(mn[d]);
      mx[d] = wave_max_f32(mx[d]);
    }
    if (l == 0) {
      for (int d = 0; d < 3; ++d) {
        nodes[t].xmin[d] = mn[d];
        nodes[t].xmax[d] = mx[d];
      }
    }
  }
  // one pass over the addends, 8 loads in flight per lane
  LaneSum ls;
  for (int i0 = 0; i0 < SM_PER; i0 += 8) {
    if ((int64_t)i0 * 64 >= cnt) break;  // uniform
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t k = (i0 + i) * 64 + l;
      v[i] = k < cnt ? addend(q, unit, X, Y, Z, M, off + k) : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) lane_sum_add(ls, v[i]);
  }
  int32_t tl;
  bool bad;
  const uint64_t tot = wave_lane_sum(ls, tl, bad);
  if (!bad) {
    if (l == 0) cms[4 * (t - lb) + q] = sd_from_units(tot, tl);
    return;
  }
  if (l == 0) {
    const int64_t k = (int64_t)atomicAdd((unsigned long long *)&ctr->cm_fail, 1ull);
    cm_list[k] = (int32_t)(4 * (t - lb) + q);
  }
}

// The small-node sums that round: the addends staged in LDS in order and
// added by sd_block (exact sequential double additions), a wave each.
__global__ void __launch_bounds__(64)
k_cm_small_ordered(const BNode *nodes, int64_t lb, bool unit, const float *X, const float *Y,
                   const float *Z, const float *M, sd_t *cms, const int32_t *cm_list,
                   const Counters *ctr) {
  __shared__ float buf[CM_SMALL];
  const int l = lane_id();
  for (int64_t k = blockIdx.x; k < ctr->cm_fail; k += gridDim.x) {
    const int32_t item = cm_list[k];
    const int64_t t = lb + item / 4;
    const int q = item % 4;
    const int64_t cnt = nodes[t].count, off = nodes[t].offset;
    sd_t S = sd_zero();
    __syncthreads();  // the previous sum has read buf
    stage_addends<8>(buf, q, unit, X, Y, Z, M, off, cnt);
    __syncthreads();
    sd_block(S, buf, (int)cnt);
    if (l == 0) cms[item] = S;
  }
}

// xc = (float)(sum w x / sum w) per dimension of a non-empty node of the
// level, from the four exact sums (BGQCM.c:208-210).
__device__ __forceinline__ void cm_finish_node(BNode &n, bool unit, const sd_t *S) {
  const int64_t cnt = n.count;
  if (cnt == 0) return;
  // unit masses: the mass sum is the count, exactly (in units of 2^0, as the
  // integer-sum path forms it from addends 1.0)
  const sd_t Sm = unit ? sd_from_units((uint64_t)cnt, 0) : S[3];
  for (int d = 0; d < 3; ++d) n.xc[d] = sd_div_to_f32(S[d], Sm);
}

// ------------------------------------------------------------ cm (big)
// Block table for nodes with count > CM_SMALL.
__global__ void __launch_bounds__(64)
k_cm_plan(BNode *nodes, int64_t lb, int64_t le, int32_t *cnode, Counters *ctr) {
  // one lane per node: a wave scan of the block counts and one atomic per
  // wave reserve the blocks; then the wave fills them node by node
  const int l = lane_id();
  const int64_t t0 = lb + (int64_t)blockIdx.x * 64;
  const int64_t t = t0 + l;
  int64_t nb = 0;
  if (t < le) {
    const int64_t cnt = nodes[t].count;
    if (cnt > CM_SMALL) nb = (cnt + CM_BLK - 1) / CM_BLK;
    else nodes[t].nblk = 0;
  }
  int64_t inc = nb;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = (int64_t)__shfl_up((long long)inc, o);
    if (l >= o) inc += y;
  }
  int64_t base = 0;
  if (l == 63 && inc > 0)
    base = (int64_t)atomicAdd((unsigned long long *)&ctr->cm_blocks, (unsigned long long)inc);
  base = (int64_t)__shfl((long long)base, 63);
  const int64_t b0 = base + inc - nb;
  if (nb > 0) {
    nodes[t].blk0 = b0;
    nodes[t].nblk = nb;
  }
  // a few blocks by the node's own lane, longer tables by the whole wave
  if (nb > 0 && nb <= 4)
    for (int64_t b = 0; b < nb; ++b) cnode[b0 + b] = (int32_t)t;
  for (uint64_t big = __ballot(nb > 4); big; big &= big - 1) {
    const int j = __builtin_ctzll(big);
    const int64_t nbj = (int64_t)__shfl((long long)nb, j);
    const int64_t b0j = (int64_t)__shfl((long long)b0, j);
    for (int64_t b = l; b < nbj; b += 64) cnode[b0j + b] = (int32_t)(t0 + j);
  }
}

// Approximate sums and bbox per block (one wave per block).  With unit
// masses (UNIT) the mass sum's record is written directly: approximate sum
// and integer sum len in units of 2^0.  The block's integer sum of a
// nonnegative sum is taken in units of its finest lsb (LaneSum); one that
// reaches 2^53 units is marked as not an integer there, which leaves its
// node off the integer path just as the node's total would.
template <bool UNIT>
__global__ void __launch_bounds__(64)
k_cm_blk_stats(const BNode *nodes, const int32_t *cnode, CmBlock *blk, const Counters *ctr,
               const float *X, const float *Y, const float *Z, const float *M) {
  constexpr int NQ = UNIT ? 3 : 4;  // sums formed from the addends
  constexpr int RU = 4;             // elements per lane per load round
  const int l = lane_id();
  for (int64_t b = blockIdx.x; b < ctr->cm_blocks; b += gridDim.x) {
    CmBlock &B = blk[b];
    int64_t start;
    int len;
    blk_range(nodes[cnode[b]], b, CM_BLK, start, len);
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    LaneSum ls[4];
    bool neg[4] = {false, false, false, false};
    // rounds of RU elements per lane: the loads of a round are in flight
    // together and few values stay live
#pragma unroll 1
    for (int i0 = 0; i0 < CM_PER; i0 += RU) {
      float xv[RU], yv[RU], zv[RU], wv[RU];
#pragma unroll
      for (int i = 0; i < RU; ++i) {
        const int k = (i0 + i) * 64 + l;
        const int64_t e = start + (k < len ? k : 0);
        wv[i] = UNIT ? 1.f : M[e];
        xv[i] = X[e];
        yv[i] = Y[e];
        zv[i] = Z[e];
      }
#pragma unroll
      for (int i = 0; i < RU; ++i) {
        const bool v = (i0 + i) * 64 + l < len;
        mn[0] = fminf(mn[0], v ? xv[i] : INFINITY); mx[0] = fmaxf(mx[0], v ? xv[i] : -INFINITY);
        mn[1] = fminf(mn[1], v ? yv[i] : INFINITY); mx[1] = fmaxf(mx[1], v ? yv[i] : -INFINITY);
        mn[2] = fminf(mn[2], v ? zv[i] : INFINITY); mx[2] = fmaxf(mx[2], v ? zv[i] : -INFINITY);
        const float w = wv[i];
        // the addends of BGQCM.c (w * x is x exactly when w = 1); zero outside
        const float a[4] = {v ? (UNIT ? xv[i] : w * xv[i]) : 0.f, v ? (UNIT ? yv[i] : w * yv[i]) : 0.f,
                            v ? (UNIT ? zv[i] : w * zv[i]) : 0.f, v ? w : 0.f};
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          s[q] += a[q];
          neg[q] = neg[q] || a[q] < 0.f;
          lane_sum_add(ls[q], a[q]);
        }
      }
    }
    int32_t lsb[4] = {1 << 20, 1 << 20, 1 << 20, 1 << 20};
    uint64_t un[4] = {0, 0, 0, 0};
    bool bad[4] = {false, false, false, false};
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      s[q] = wave_sum_f32(s[q]);
      un[q] = wave_lane_sum(ls[q], lsb[q], bad[q]);
      // a negative addend: no integer path for this sum (INT32_MIN)
      if (__any(neg[q])) lsb[q] = INT32_MIN;
      bad[q] = bad[q] || lsb[q] == INT32_MIN;
    }
    if (UNIT) {
      // len ones: the float sum, the lsb of 1.0 and the count are exact
      s[3] = (float)len;
      lsb[3] = 0;
      un[3] = (uint64_t)len;
      bad[3] = false;
    }
    for (int d = 0; d < 3; ++d) {
      mn[d] = wave_min_f32(mn[d]);
      mx[d] = wave_max_f32(mx[d]);
    }
    if (l == 0) {
      for (int q = 0; q < 4; ++q) {
        B.approx[q] = s[q];
        B.lsb[q] = lsb[q];
        B.units[q] = bad[q] ? ~0ull : un[q];
      }
      for (int d = 0; d < 3; ++d) {
        B.bmin[d] = mn[d];
        B.bmax[d] = mx[d];
      }
    }
  }
}

// Big nodes' reductions of their block records in two passes: a wave per
// 64-block chunk, then a wave per node over its chunks (the top levels hold
// one to a few nodes of up to ~18k blocks, so no pass runs a whole node on
// one compute unit).  They give the exclusive prefix of the approximate
// block sums (only a guess of the running sum's binade, so any summation
// order serves), the node bbox (min/max: order-free), and the integer-sum
// path: with t the finest lsb of a sum's addends over the node, every addend
// a nonnegative multiple of 2^t and their total below 2^(t+53), no partial
// sum of the sequential double additions rounds (each is such a multiple
// below 2^(t+53)), so the sum is that integer times 2^t: written to cms and
// flagged done (bit q of cmflags) for the ordered path to skip.
__device__ __forceinline__ uint64_t units_add_sat(uint64_t s, uint64_t ub, int32_t tb, int32_t tq) {
  if (s >= SD_M53 || ub == ~0ull) return SD_M53;
  if (ub == 0) return s;
  const int32_t sh = tb - tq;  // >= 0
  if (sh >= 53 || (ub >> (53 - sh)) != 0) return SD_M53;
  const uint64_t v = s + (ub << sh);
  return v >= SD_M53 ? SD_M53 : v;
}

constexpr int PFX_WAVES = 16;  // waves per workgroup of the top levels' per-node scans
constexpr int PFX_T = 64 * PFX_WAVES;

// One wave per chunk (the up to 64 blocks from a leader block, the first of
// every 64 of a node with more than 64 blocks), a lane per block: the local exclusive prefix of the
// approximate sums (into the block's ein; k_cm_blk_maps adds the chunk's
// base), and the chunk's totals, bbox, finest lsb and integer sum in units
// of that lsb (saturating at 2^53: a chunk at the bound leaves its node
// above it in any finer unit), kept in the leader.
__global__ void __launch_bounds__(64)
k_cm_chunk(const BNode *nodes, const int32_t *cnode, CmBlock *blk, const Counters *ctr) {
  const int l = lane_id();
  const int64_t nblocks = ctr->cm_blocks;
  // a wave per 64-block range of the table, lane l testing block r + l; the
  // range's leaders are taken in turn (one at the top levels, where leaders
  // sit at multiples of 64)
  for (int64_t r = 64 * (int64_t)blockIdx.x; r < nblocks; r += 64 * (int64_t)gridDim.x) {
    uint64_t lead;
    {
      const int64_t c = r + l;
      bool isl = false;
      if (c < nblocks) {
        const BNode &nd = nodes[cnode[c]];
        isl = ((c - nd.blk0) & 63) == 0 && nd.nblk > 64;  // k_cm_node takes single chunks
      }
      lead = __ballot(isl);
    }
    while (lead) {
      const int64_t c = r + __builtin_ctzll(lead);
      lead &= lead - 1;
      const BNode &nd = nodes[cnode[c]];
      const int64_t rel = c - nd.blk0;
      const int nbc = (int)min<int64_t>(64, nd.nblk - rel);
      const bool v = l < nbc;
      CmBlock &B = blk[c + (v ? l : 0)];
      float a[4], mn[3], mx[3];
      int32_t lsb[4];
      uint64_t un[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = v ? B.approx[q] : 0.f;
        lsb[q] = v ? B.lsb[q] : (1 << 20);
        un[q] = v ? B.units[q] : 0ull;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        mn[d] = v ? B.bmin[d] : INFINITY;
        mx[d] = v ? B.bmax[d] : -INFINITY;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float inc = a[q];  // inclusive scan over the lanes
        for (int o = 1; o < 64; o <<= 1) {
          const float y = __shfl_up(inc, o);
          if (l >= o) inc += y;
        }
        if (v) B.ein[q] = inc - a[q];
        a[q] = __shfl(inc, 63);  // the chunk's total
        const int32_t tc = wave_min_i32(lsb[q]);
        uint64_t su = 0;
        if (tc != INT32_MIN) su = units_add_sat(0ull, un[q], lsb[q], tc);
        su = wave_sum_u64(su);  // 64 values of at most 2^53: no wrap
        lsb[q] = tc;
        un[q] = su >= SD_M53 ? SD_M53 : su;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        mn[d] = wave_min_f32(mn[d]);
        mx[d] = wave_max_f32(mx[d]);
      }
      if (l == 0) {
        CmBlock &C = blk[c];
        for (int q = 0; q < 4; ++q) {
          C.capprox[q] = a[q];
          C.clsb[q] = lsb[q];
          C.cunits[q] = lsb[q] == INT32_MIN ? SD_M53 : un[q];
        }
        for (int d = 0; d < 3; ++d) {
          C.cbmin[d] = mn[d];
          C.cbmax[d] = mx[d];
        }
      }
    }
  }
}

// One wave per big node over its chunks: the chunks' exclusive prefix
// (chunk base), the node bbox, the node's finest lsb tq and the integer sum
// in units of 2^tq (saturating).
__global__ void __launch_bounds__(64)
k_cm_node(BNode *nodes, int64_t lb, int64_t le, CmBlock *blk, int32_t *cmflags, sd_t *cms) {
  const int l = lane_id();
  for (int64_t t = lb + blockIdx.x; t < le; t += gridDim.x) {
    if (nodes[t].count <= CM_SMALL) continue;
    const int64_t b0 = nodes[t].blk0, nc = (nodes[t].nblk + 63) / 64;
    float base[4] = {0.f, 0.f, 0.f, 0.f};
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    int32_t lsb[4] = {1 << 20, 1 << 20, 1 << 20, 1 << 20};
    if (nc == 1) {
      // at most 64 blocks: their records directly, a lane per block (the
      // node is its only chunk, whose base is zero)
      const int nb = (int)nodes[t].nblk;
      const bool v = l < nb;
      CmBlock &B = blk[b0 + (v ? l : 0)];
      uint64_t su[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float a = v ? B.approx[q] : 0.f;
        float inc = a;
        for (int o = 1; o < 64; o <<= 1) {
          const float y = __shfl_up(inc, o);
          if (l >= o) inc += y;
        }
        if (v) B.ein[q] = inc - a;
        const int32_t lq = v ? B.lsb[q] : (1 << 20);
        const int32_t tq = wave_min_i32(lq);
        uint64_t u = 0;
        if (tq != INT32_MIN && v) u = units_add_sat(0ull, B.units[q], lq, tq);
        su[q] = wave_sum_u64(u);  // 64 values of at most 2^53
        lsb[q] = tq;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        mn[d] = wave_min_f32(v ? B.bmin[d] : INFINITY);
        mx[d] = wave_max_f32(v ? B.bmax[d] : -INFINITY);
      }
      if (l == 0) {
        for (int q = 0; q < 4; ++q) blk[b0].cbase[q] = 0.f;
        for (int d = 0; d < 3; ++d) {
          nodes[t].xmin[d] = mn[d];
          nodes[t].xmax[d] = mx[d];
        }
        int32_t f = 0;
        for (int q = 0; q < 4; ++q) {
          // a negative addend, all addends zero, or a total that rounds
          if (lsb[q] == INT32_MIN || lsb[q] == (1 << 20) || su[q] >= SD_M53) continue;
          cms[4 * (t - lb) + q] = sd_from_units(su[q], lsb[q]);
          f |= 1 << q;
        }
        cmflags[t - lb] = f;
      }
      continue;
    }
    for (int64_t k0 = 0; k0 < nc; k0 += 64) {
      const bool v = k0 + l < nc;
      CmBlock &C = blk[b0 + 64 * (v ? k0 + l : 0)];  // the chunk's leader
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float a = v ? C.capprox[q] : 0.f;
        float inc = a;
        for (int o = 1; o < 64; o <<= 1) {
          const float y = __shfl_up(inc, o);
          if (l >= o) inc += y;
        }
        if (v) C.cbase[q] = base[q] + (inc - a);
        base[q] += __shfl(inc, 63);
        if (v) lsb[q] = min(lsb[q], C.clsb[q]);
      }
      if (v)
        for (int d = 0; d < 3; ++d) {
          mn[d] = fminf(mn[d], C.cbmin[d]);
          mx[d] = fmaxf(mx[d], C.cbmax[d]);
        }
    }
    int32_t tq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) tq[q] = wave_min_i32(lsb[q]);
    uint64_t su[4] = {0, 0, 0, 0};
    for (int64_t k = l; k < nc; k += 64) {
      const CmBlock &C = blk[b0 + 64 * k];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (tq[q] != INT32_MIN) su[q] = units_add_sat(su[q], C.cunits[q], C.clsb[q], tq[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) su[q] = wave_sum_u64(su[q]);  // 64 values of at most 2^53
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      mn[d] = wave_min_f32(mn[d]);
      mx[d] = wave_max_f32(mx[d]);
    }
    if (l == 0) {
      for (int d = 0; d < 3; ++d) {
        nodes[t].xmin[d] = mn[d];
        nodes[t].xmax[d] = mx[d];
      }
      int32_t f = 0;
      for (int q = 0; q < 4; ++q) {
        // a negative addend, all addends zero, or a total that rounds
        if (tq[q] == INT32_MIN || tq[q] == (1 << 20) || su[q] >= SD_M53) continue;
        cms[4 * (t - lb) + q] = sd_from_units(su[q], tq[q]);
        f |= 1 << q;
      }
      cmflags[t - lb] = f;
    }
  }
}

__device__ __forceinline__ int32_t grid_exp_guess(float ein) {
  uint32_t u = sd_f2u(ein);
  int32_t ef = (int32_t)((u >> 23) & 255u);
  if (ein <= 0.f || ef == 0 || ef == 255) return INT32_MIN;
  return (ef - 127) - 52;
}


// Fixed-grid maps of each block for its guessed binade: the block is staged
// in LDS, each lane composes CM_PER consecutive addends in order, then the 64
// lane maps are composed in lane order.
__global__ void __launch_bounds__(64)
k_cm_blk_maps(const BNode *nodes, const int32_t *cnode, CmBlock *blk, const Counters *ctr,
              bool unit, const float *X, const float *Y, const float *Z, const float *M,
              int64_t lb, const int32_t *cmflags) {
  __shared__ float buf[64 * (CM_PER + 1)];
  const int q = blockIdx.y;  // one wave per (block, sum)
  for (int64_t b = blockIdx.x; b < ctr->cm_blocks; b += gridDim.x) {
    CmBlock &B = blk[b];
    const int64_t t = cnode[b];
    if (cmflags[t - lb] & (1 << q)) continue;  // summed exactly
    const int l = lane_id();
    int64_t start;
    int len;
    blk_range(nodes[t], b, CM_BLK, start, len);
    const int64_t b0 = nodes[t].blk0;
    const float ein = B.ein[q] + blk[b0 + ((b - b0) & ~(int64_t)63)].cbase[q];
    const int32_t E = grid_exp_guess(ein);
    sd_map acc = sd_map_identity();
    if (E == INT32_MIN) {
      acc.bad = 1;
    } else {
      __syncthreads();  // the previous block's maps have read buf
#pragma unroll
      for (int k = l; k < CM_BLK; k += 64) {
        float v = k < len ? addend(q, unit, X, Y, Z, M, start + k) : 0.f;
        buf[(k / CM_PER) * (CM_PER + 1) + (k % CM_PER)] = v;
      }
      __syncthreads();
      for (int k = 0; k < CM_PER; ++k)
        acc = sd_map_compose(acc, sd_map_of(buf[l * (CM_PER + 1) + k], E));
      acc = wave_compose_maps(acc);
    }
    if (l == 0) {
      B.d0[q] = acc.d0;
      B.d1[q] = acc.d1;
      B.bad[q] = acc.bad;
      B.E[q] = E;
    }
  }
}

// Window maps: the first block of every 64-block window of a node composes
// the window's block maps in order when they share one grid (one wave per
// block; waves of other blocks return at once).
__global__ void __launch_bounds__(64)
k_cm_win_maps(const BNode *nodes, const int32_t *cnode, CmBlock *blk, const Counters *ctr) {
  for (int64_t b = blockIdx.x; b < ctr->cm_blocks; b += gridDim.x) {
    const int64_t t = cnode[b];
    const int64_t b0 = nodes[t].blk0, nb = nodes[t].nblk;
    if ((b - b0) % 64 != 0) continue;
    const int l = lane_id();
    const int cnt = (int)min<int64_t>(64, b0 + nb - b);
    for (int q = 0; q < 4; ++q) {
      sd_map m = sd_map_identity();
      int32_t E = 0;
      int bad = 1;
      if (l < cnt) {
        const CmBlock &B = blk[b + l];
        m.d0 = B.d0[q];
        m.d1 = B.d1[q];
        E = B.E[q];
        bad = B.bad[q];
      }
      const int32_t E0 = readlane_i32(E, 0);
      const int ok = !__any(l < cnt && (bad || E != E0));
      if (l >= cnt) m = sd_map_identity();
      m = wave_compose_maps(m);
      if (l == 0) {
        blk[b].wd0[q] = m.d0;
        blk[b].wd1[q] = m.d1;
        blk[b].wE[q] = E0;
        blk[b].wok[q] = ok && !m.bad;
      }
    }
  }
}

// Chain the block maps of each big node in order (64 blocks' maps are read
// at once, one per lane, then applied serially); a block whose map does not
// apply (binade guess wrong or binade crossed) is staged in LDS and
// re-accumulated chunk by chunk.
__global__ void __launch_bounds__(64)
k_cm_blk_chain(BNode *nodes, int64_t lb, int64_t le, bool unit, const CmBlock *blk,
               const float *X, const float *Y, const float *Z,
               const float *M, Counters *ctr, sd_t *cms, const int32_t *cmflags) {
  __shared__ float buf[CM_BLK];
  const int q = blockIdx.y;  // one wave per (node, sum): x, y, z, mass
  const int l = lane_id();
  for (int64_t t = lb + blockIdx.x; t < le; t += gridDim.x) {
    if (nodes[t].count <= CM_SMALL) continue;
    if (cmflags[t - lb] & (1 << q)) continue;  // summed exactly
    const int64_t b0 = nodes[t].blk0, nb = nodes[t].nblk;
    sd_t Sq = sd_zero();
    // the window maps of 64 windows at once, one per lane (the loop over the
    // windows is serial; it must not wait on a load per window)
    uint64_t wd0 = 0, wd1 = 0;
    int32_t wok = 0, wE = 0;
    for (int64_t bb = 0; bb < nb; bb += 64) {
      const int wi = (int)((bb >> 6) & 63);
      if (wi == 0) {
        const int64_t lb0 = bb + 64 * (int64_t)l;  // leader block of lane l's window
        wok = 0;
        if (lb0 < nb) {
          const CmBlock &W = blk[b0 + lb0];
          wd0 = W.wd0[q];
          wd1 = W.wd1[q];
          wok = W.wok[q];
          wE = W.wE[q];
        }
      }
      // the whole window at once when its blocks share a grid and S stays in it
      {
        sd_map g;
        g.d0 = readlane_u64(wd0, wi);
        g.d1 = readlane_u64(wd1, wi);
        g.bad = !readlane_i32(wok, wi);
        if (sd_map_apply(&Sq, g, readlane_i32(wE, wi))) continue;
      }
      uint64_t md0 = 0, md1 = 0;
      int mbad = 1, mE = 0;
      if (bb + l < nb) {
        const CmBlock &B = blk[b0 + bb + l];
        md0 = B.d0[q];
        md1 = B.d1[q];
        mbad = B.bad[q];
        mE = B.E[q];
      }
      const int cnt = (int)min<int64_t>(64, nb - bb);
      {
        for (int i = 0; i < cnt;) {  // uniform
          if (Sq.m != 0 && !Sq.s) {
            // ordered scan of the block maps from i built for S's binade; the
            // longest prefix that keeps S in the binade is applied at once
            const int32_t E = Sq.e;
            sd_map m = sd_map_identity();
            if (l >= i && l < cnt) {
              m.d0 = md0;
              m.d1 = md1;
              m.bad = mbad || mE != E;
            }
            for (int o = 1; o < 64; o <<= 1) {
              sd_map h;
              h.d0 = (uint64_t)__shfl_up((unsigned long long)m.d0, o);
              h.d1 = (uint64_t)__shfl_up((unsigned long long)m.d1, o);
              h.bad = __shfl_up(m.bad, o);
              if (l >= o) m = sd_map_compose(h, m);
            }
            sd_t S2 = Sq;
            const bool ok = l >= i && l < cnt && sd_map_apply(&S2, m, E);
            const uint64_t bal = ((uint64_t)__ballot(ok) >> i) | (cnt - i < 64 ? ~0ull << (cnt - i) : 0ull);
            const int j = min(cnt, i + ((~bal == 0ull) ? 64 - i : (int)__builtin_ctzll(~bal)));
#ifdef HSR_CM_STATS
            if (l == 0) atomicAdd((unsigned long long *)&ctr->pad[0], 1ull);
#endif
            if (j > i) {
              Sq.m = readlane_u64(S2.m, j - 1);
              Sq.e = readlane_i32(S2.e, j - 1);
              Sq.s = readlane_i32(S2.s, j - 1);
              i = j;
              continue;
            }
          }
          // block i does not fit one grid: its addends, exactly, from LDS
#ifdef HSR_CM_STATS
          if (l == 0) atomicAdd((unsigned long long *)&ctr->pad[1], 1ull);
#endif
          int64_t start;
          int len;
          blk_range(nodes[t], b0 + bb + i, CM_BLK, start, len);
          __syncthreads();
          stage_addends<8>(buf, q, unit, X, Y, Z, M, start, len);
          __syncthreads();
          sd_block(Sq, buf, (int)len);
          ++i;
        }
      }
    }
    if (l == 0) cms[4 * (t - lb) + q] = Sq;
  }
}

// ------------------------------------------------------------ classify
// Leaves (count <= nd) get tdr and, for count > 1, ppm = sequential float
// sum of member masses (pp with TDPTS = 1: mass*(1+0+0)); other nodes get
// the split dimension (longest edge, ties x > y > z as coded at
// RCBForceTree.cxx:843-852).  Returns whether the node splits.
__device__ __forceinline__ bool classify_node(BNode &n, int64_t nd, float ppc, bool unit_ppm) {
  n.cl = n.cr = -1;
  if (n.count <= nd) {
    n.kind = NODE_LEAF;
    if (n.count == 0) {
      n.xc[0] = n.xc[1] = n.xc[2] = __int_as_float(0x7fc00000);
    }
    n.tdr = ppc * pptdr_dev(n.xmin, n.xmax, n.xc);
    // leaves with more than one member: k_leaf_ppm, or with unit masses
    // (unit_ppm) the in-order float sum of count ones: the count, exact up
    // to 2^24, where adding 1 rounds back to 2^24 (ties to even)
    n.ppm = (unit_ppm && n.count > 1) ? (n.count < (1 << 24) ? (float)n.count : 16777216.f) : 0.f;
    return false;
  }
  n.kind = NODE_SPLIT;
  float xl0 = n.xmax[0] - n.xmin[0];
  float xl1 = n.xmax[1] - n.xmin[1];
  float xl2 = n.xmax[2] - n.xmin[2];
  int d;
  if (xl0 > xl1 && xl0 > xl2) d = 0;
  else if (xl1 > xl2) d = 1;
  else d = 2;
  n.split_d = d;
  return true;
}

// Lowest set bit of a finite nonzero float, as a power-of-two exponent.
__device__ __forceinline__ int f32_lsb_exp(uint32_t b) {
  const int e = (int)((b >> 23) & 0xffu);
  const uint32_t mant = b & 0x7fffffu;
  return e == 0 ? -149 + __builtin_ctz(mant) : e - 150 + __builtin_ctz(mant | 0x800000u);
}

// Leaf pseudo mass (pp with TDPTS = 1: the members' masses times
// (1 + 0 + 0), summed in member order in float; RCBForceTree.cxx:536-569,
// 787-797), one wave per leaf with more than one member.  When every mass is
// a multiple of the finest lsb u among them and their magnitudes sum below
// 2^24 u, every partial sum of the sequential order is exact, so that sum is
// the exact integer sum in units of u, which the wave forms in any order
// (unit masses always take this path).  Otherwise lane 0 adds them in order.
__global__ void __launch_bounds__(64)
k_leaf_ppm(BNode *nodes, int64_t lb, int64_t le, int64_t nd, const float *M) {
  const int l = lane_id();
  for (int64_t t = lb + blockIdx.x; t < le; t += gridDim.x) {
    BNode &n = nodes[t];
    const int64_t cnt = n.count, off = n.offset;
    if (cnt > nd || cnt <= 1) continue;
    int emin = 1 << 20;
    bool bad = false;
    for (int64_t i = l; i < cnt; i += 64) {
      const float m = M[off + i] * 1.0f;
      const uint32_t b = __float_as_uint(m);
      if (((b >> 23) & 0xffu) == 0xffu) bad = true;  // inf or nan
      else if ((b & 0x7fffffffu) != 0u) emin = min(emin, f32_lsb_exp(b));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emin = min(emin, __shfl_xor(emin, o));
    int64_t sum = 0;
    uint64_t mag = 0;
    for (int64_t i = l; i < cnt && !bad; i += 64) {
      const uint32_t b = __float_as_uint(M[off + i] * 1.0f);
      if ((b & 0x7fffffffu) == 0u) continue;
      const int e = (int)((b >> 23) & 0xffu);
      const uint32_t mant = e == 0 ? (b & 0x7fffffu) : ((b & 0x7fffffu) | 0x800000u);
      const int sh = f32_lsb_exp(b) - emin;  // |m| = odd * 2^(lsb), odd = mant >> ctz
      if (sh > 32) {
        bad = true;
        break;
      }
      const uint64_t v = (uint64_t)(mant >> __builtin_ctz(mant)) << sh;
      mag += v;
      sum += (b >> 31) ? -(int64_t)v : (int64_t)v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sum += __shfl_xor(sum, o);
      mag += __shfl_xor(mag, o);
      bad = __shfl_xor((int)bad, o) || bad;
    }
    if (l != 0) continue;
    if (!bad && mag < (1ull << 24)) {
      n.ppm = sum == 0 ? 0.f : ldexpf((float)sum, emin);
    } else {
      float ppm = 0.f;
      for (int64_t i = 0; i < cnt; ++i) ppm += M[off + i] * 1.0f;
      n.ppm = ppm;
    }
  }
}

// Copy the final permutation of nodes that stop here (leaves at classify
// time read buffer `P`; degenerate splits read the partition's output).
__global__ void __launch_bounds__(64)
k_finalize_leaves(const BNode *__restrict__ nodes, int64_t lb, int64_t le, int kind, bool ident,
                  const int32_t *__restrict__ P, int32_t *__restrict__ Pfin, bool unit,
                  const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
                  const float *__restrict__ M, float *__restrict__ xs, float *__restrict__ ys,
                  float *__restrict__ zs, float *__restrict__ ms) {
  // the node's elements are in their final positions: their input index
  // and their sorted x, y, z, m (read in order here rather than gathered
  // through the permutation afterwards)
  for (int64_t t = lb + blockIdx.x; t < le; t += gridDim.x) {
    if (nodes[t].kind != kind) continue;
    int64_t off = nodes[t].offset, cnt = nodes[t].count;
    for (int64_t k = lane_id(); k < cnt; k += 64) {
      const int64_t e = off + k;
      Pfin[e] = ident ? (int32_t)e : P[e];
      xs[e] = X[e];
      ys[e] = Y[e];
      zs[e] = Z[e];
      ms[e] = unit ? 1.f : M[e];
    }
  }
}

// The particles' j-records from the sorted arrays.
__global__ void k_pj_particles(int64_t n, const float *xs, const float *ys, const float *zs,
                               const float *ms, PJ *pj) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) pj[k] = PJ{xs[k], ys[k], zs[k], ms[k]};
}

// ------------------------------------------------------------ partition
__global__ void __launch_bounds__(64)
k_part_plan(BNode *nodes, int64_t lb, int64_t le, bool unit, const sd_t *cms, int64_t nd,
            float ppc, bool unit_ppm, int32_t *pnode, Counters *ctr) {
  // one lane per node: its centre of mass and class (cm_finish_node,
  // classify_node), then a wave scan of the block counts of the splitting
  // nodes and one atomic per wave reserve their partition blocks; then the
  // wave fills them node by node
  const int l = lane_id();
  const int64_t t0 = lb + (int64_t)blockIdx.x * 64;
  const int64_t t = t0 + l;
  int64_t nb = 0;
  if (t < le) {
    BNode &n = nodes[t];
    cm_finish_node(n, unit, cms + 4 * (t - lb));
    if (classify_node(n, nd, ppc, unit_ppm)) nb = (n.count + PB - 1) / PB;
  }
  int64_t inc = nb;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = (int64_t)__shfl_up((long long)inc, o);
    if (l >= o) inc += y;
  }
  int64_t base = 0;
  if (l == 63 && inc > 0)
    base = (int64_t)atomicAdd((unsigned long long *)&ctr->part_blocks, (unsigned long long)inc);
  base = (int64_t)__shfl((long long)base, 63);
  const int64_t b0 = base + inc - nb;
  if (nb > 0) {
    nodes[t].blk0 = b0;
    nodes[t].nblk = nb;
  }
  // a few blocks by the node's own lane, longer tables by the whole wave
  if (nb > 0 && nb <= 4)
    for (int64_t b = 0; b < nb; ++b) pnode[b0 + b] = (int32_t)t;
  for (uint64_t big = __ballot(nb > 4); big; big &= big - 1) {
    const int j = __builtin_ctzll(big);
    const int64_t nbj = (int64_t)__shfl((long long)nb, j);
    const int64_t b0j = (int64_t)__shfl((long long)b0, j);
    for (int64_t b = l; b < nbj; b += 64) pnode[b0j + b] = (int32_t)(t0 + j);
  }
}

// Keys < pivot per block (RCBForceTree.cxx:638-645), pivot = xc[d].
__global__ void __launch_bounds__(256)
k_part_count(const BNode *__restrict__ nodes, const int32_t *__restrict__ pnode,
             PartBlock *__restrict__ pb, const Counters *__restrict__ ctr, const float *__restrict__ X,
             const float *__restrict__ Y, const float *__restrict__ Z, uint64_t *__restrict__ mask) {
  for (int64_t b = blockIdx.x; b < ctr->part_blocks; b += gridDim.x) {
    const BNode &n = nodes[pnode[b]];
    int64_t bstart;
    int blen;
    blk_range(n, b, PB, bstart, blen);
    const float *key = coord(n.split_d, X, Y, Z);
    float pv = n.xc[n.split_d];
    int c = 0;
    // the flags also go to the block's mask words (64 elements each), which
    // the rank kernel reads instead of the keys; the keys are all loaded
    // first (restrict-qualified buffers), so their loads are in flight together
    float kv[PB / 256];
#pragma unroll
    for (int r = 0; r < PB / 256; ++r) {
      const int k = r * 256 + (int)threadIdx.x;
      kv[r] = key[bstart + (k < blen ? k : 0)];
    }
#pragma unroll
    for (int r = 0; r < PB / 256; ++r) {
      const int k = r * 256 + (int)threadIdx.x;
      const bool f = k < blen && kv[r] < pv;
      const uint64_t bal = __ballot(f);
      if (lane_id() == 0) mask[b * (PB / 64) + (k >> 6)] = bal;
      c += f;
    }
    __shared__ int red[4];
    __syncthreads();  // thread 0 has read the previous block's red
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) pb[b].cnt = red[0] + red[1] + red[2] + red[3];
  }
}

// In-order exclusive scan of block counts per node, a workgroup of NW waves
// per node (16 at the top levels, where a node holds thousands of blocks,
// one below); degenerate splits (is == 0 or is == count,
// RCBForceTree.cxx:727-729) are marked.
template <int NW>
__global__ void __launch_bounds__(64 * NW)
k_part_scan(BNode *nodes, int64_t lb, int64_t le, PartBlock *pb, Counters *ctr) {
  constexpr int T = 64 * NW;
  __shared__ int64_t wsum[NW];
  const int l = lane_id(), w = threadIdx.x >> 6;
  for (int64_t t = lb + blockIdx.x; t < le; t += gridDim.x) {
    if (nodes[t].kind != NODE_SPLIT) continue;  // uniform over the workgroup
    const int64_t b0 = nodes[t].blk0, nb = nodes[t].nblk;
    int64_t run = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += T) {
      const int64_t b = c0 + threadIdx.x;
      const int64_t v = b < nb ? pb[b0 + b].cnt : 0;
      int64_t s = v;  // inclusive scan over the wave
      for (int o = 1; o < 64; o <<= 1) {
        const int64_t u = (int64_t)__shfl_up((long long)s, o);
        if (l >= o) s += u;
      }
      if (l == 63) wsum[w] = s;
      __syncthreads();
      int64_t before = 0, all = 0;
      for (int k = 0; k < NW; ++k) {
        if (k < w) before += wsum[k];
        all += wsum[k];
      }
      if (b < nb) pb[b0 + b].base = run + before + s - v;
      run += all;
      __syncthreads();  // wsum is rewritten by the next chunk
    }
    if (threadIdx.x == 0) {
      nodes[t].is = run;
      if (run == 0 || run == nodes[t].count) {
        nodes[t].kind = NODE_DEGEN;
        atomicAdd((unsigned long long *)&ctr->ndegen, 1ull);
      }
    }
  }
}

// Ranks: a key < pivot at local position p moves to its rank r and is
// recorded as idx[r] = p; others at p >= is stay; others below is are
// resolved by the swap chain (dest = -1 here).
__global__ void __launch_bounds__(256)
k_part_rank(const BNode *__restrict__ nodes, const int32_t *__restrict__ pnode,
            const PartBlock *__restrict__ pb, const Counters *__restrict__ ctr,
            const uint64_t *__restrict__ mask, int32_t *__restrict__ idx, int32_t *__restrict__ rights,
            int32_t *__restrict__ dest) {
  // the block's 16 mask words (set by k_part_count, zero past the block)
  // are read at once and their counts shared through LDS with one barrier
  __shared__ int wcnt[PB / 64];
  const int w = threadIdx.x >> 6, l = lane_id();
  for (int64_t b = blockIdx.x; b < ctr->part_blocks; b += gridDim.x) {
    const BNode &n = nodes[pnode[b]];
    int64_t bstart;
    int blen;
    blk_range(n, b, PB, bstart, blen);
    const int64_t off = n.offset, is = n.is, base = pb[b].base;
    uint64_t bal[PB / 256];
#pragma unroll
    for (int r = 0; r < PB / 256; ++r) bal[r] = mask[b * (PB / 64) + r * 4 + w];
    if (l == 0)
#pragma unroll
      for (int r = 0; r < PB / 256; ++r) wcnt[r * 4 + w] = __popcll(bal[r]);
    __syncthreads();
    int64_t run = base;
#pragma unroll
    for (int r = 0; r < PB / 256; ++r) {
      const int k = r * 256 + (int)threadIdx.x;
      int wbase = 0, tot = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int v = wcnt[r * 4 + i];
        wbase += i < w ? v : 0;
        tot += v;
      }
      if (k < blen) {
        const int64_t e = bstart + k, p = e - off;
        const int64_t lefts_before = run + wbase + __popcll(bal[r] & lanemask_lt());
        if ((bal[r] >> l) & 1ull) {
          idx[off + lefts_before] = (int32_t)p;
          dest[e] = (int32_t)lefts_before;
        } else {
          rights[off + (p - lefts_before)] = (int32_t)p;
          dest[e] = p >= is ? (int32_t)p : -1;
        }
      }
      run += tot;
    }
    __syncthreads();  // wcnt is rewritten for the next block
  }
}

// Swap chains (RCBForceTree.cxx:647-669): an element with key >= pivot at
// p < is ends at the first idx-iterate p -> idx[p] -> ... that is >= is.
// Between the k-th and (k+1)-th key >= pivot (positions rights[k-1] and
// rights[k]) every position is a left one, so idx[j] = j + k there: a chain
// that entered such a stretch with step k continues f, f+k, f+2k, ... up to
// rights[k]; the walk therefore jumps whole stretches arithmetically.
__global__ void __launch_bounds__(256)
k_part_chain(const BNode *__restrict__ nodes, const int32_t *__restrict__ pnode,
             const Counters *ctrc, const int32_t *__restrict__ idx,
             const int32_t *__restrict__ rights, int32_t *__restrict__ dest, Counters *ctr, int maxhop) {
  for (int64_t b = blockIdx.x; b < ctrc->part_blocks; b += gridDim.x) {
    const BNode &n = nodes[pnode[b]];
    int64_t bstart;
    int blen;
    blk_range(n, b, PB, bstart, blen);
    // positions within a node are below 2^31 (int32 permutation), so the
    // stretch arithmetic runs in 32 bits (a 64-bit division is a long
    // software sequence)
    const int64_t off = n.offset;
    const uint32_t is = (uint32_t)n.is, cnt = (uint32_t)n.count, nr = cnt - is;
    const int32_t *idx_n = idx + off;
    const int32_t *rights_n = rights + off;
    // The chains start at the keys >= pivot below is: the first entries of
    // the node's sorted rights list.  Block B walks the rights with list
    // indices in its own element range, so every active lane holds a chain
    // (most elements start none) and neighbouring lanes start at nearby
    // positions, with similar chain lengths.
    const int64_t r0 = bstart - off;
    for (int k = threadIdx.x; k < blen; k += 256) {
      const int64_t r = r0 + k;
      if (r >= (int64_t)nr) break;
      const uint32_t p = (uint32_t)rights_n[r];
      if (p >= is) break;  // sorted: this thread's later entries are too
      const int64_t e = off + p;
      uint32_t j = p;
      bool done = false;
      for (int it = 0; it < maxhop; ++it) {
        const uint32_t f = (uint32_t)idx_n[j];  // f > j
        if (f >= is) {
          dest[e] = (int32_t)f;
          done = true;
          break;
        }
        const uint32_t step = f - j;  // keys >= pivot below f
        const uint32_t lim = step < nr ? (uint32_t)rights_n[step] : cnt;
        // f < is < 2^31 and step < 2^31: v < 2^32
        const uint32_t v = f + ((is - f + step - 1u) / step) * step;
        if (v < lim) {
          dest[e] = (int32_t)v;
          done = true;
          break;
        }
        j = f + ((lim - 1u - f) / step) * step;
      }
      if (!done) {
        dest[e] = -2 - (int32_t)j;  // unresolved; continue from j
        atomicAdd((unsigned long long *)&ctr->unresolved, 1ull);
      }
    }
  }
}

// Pointer-jumping fallback for long chains: jmp[off+j] = idx-iterate of j
// advanced until >= is; log2 rounds resolve every chain.
__global__ void __launch_bounds__(256)
k_part_jump_init(const BNode *nodes, const int32_t *pnode, const Counters *ctr,
                 const int32_t *idx, int32_t *jmp) {
  for (int64_t b = blockIdx.x; b < ctr->part_blocks; b += gridDim.x) {
    const BNode &n = nodes[pnode[b]];
    int64_t bstart;
    int blen;
    blk_range(n, b, PB, bstart, blen);
    for (int k = threadIdx.x; k < blen; k += 256) {
      int64_t p = bstart + k - n.offset;
      if (p < n.is) jmp[n.offset + p] = idx[n.offset + p];
    }
  }
}

__global__ void __launch_bounds__(256)
k_part_jump_round(const BNode *nodes, const int32_t *pnode,
                  const Counters *ctr, int32_t *jmp) {
  for (int64_t b = blockIdx.x; b < ctr->part_blocks; b += gridDim.x) {
    const BNode &n = nodes[pnode[b]];
    int64_t bstart;
    int blen;
    blk_range(n, b, PB, bstart, blen);
    for (int k = threadIdx.x; k < blen; k += 256) {
      int64_t p = bstart + k - n.offset;
      if (p < n.is) {
        int32_t j = jmp[n.offset + p];
        if (j < n.is) jmp[n.offset + p] = jmp[n.offset + j];
      }
    }
  }
}

__global__ void __launch_bounds__(256)
k_part_jump_final(const BNode *nodes, const int32_t *pnode,
                  const Counters *ctr, const int32_t *jmp, int32_t *dest) {
  for (int64_t b = blockIdx.x; b < ctr->part_blocks; b += gridDim.x) {
    const BNode &n = nodes[pnode[b]];
    int64_t bstart;
    int blen;
    blk_range(n, b, PB, bstart, blen);
    for (int k = threadIdx.x; k < blen; k += 256) {
      int64_t e = bstart + k;
      int32_t d = dest[e];
      if (d <= -2) dest[e] = jmp[n.offset + (-2 - d)];
    }
  }
}

__global__ void __launch_bounds__(256)
k_part_scatter(const BNode *__restrict__ nodes, const int32_t *__restrict__ pnode,
               const Counters *__restrict__ ctr, bool unit, bool ident, const int32_t *__restrict__ dest,
               const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ Z,
               const float *__restrict__ M, const int32_t *__restrict__ P, float *__restrict__ X2,
               float *__restrict__ Y2, float *__restrict__ Z2, float *__restrict__ M2,
               int32_t *__restrict__ P2) {
  // the source and destination arrays are distinct buffers: each thread's
  // PB / 256 elements are all read before any is written, so their loads
  // are in flight together
  constexpr int PER = PB / 256;
  for (int64_t b = blockIdx.x; b < ctr->part_blocks; b += gridDim.x) {
    const BNode &n = nodes[pnode[b]];
    int64_t bstart;
    int blen;
    blk_range(n, b, PB, bstart, blen);
    const int64_t off = n.offset;
    int32_t d[PER], vp[PER];
    float vx[PER], vy[PER], vz[PER], vm[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int k = u * 256 + (int)threadIdx.x;
      const int64_t e = bstart + (k < blen ? k : 0);
      d[u] = dest[e];
      vx[u] = X[e];
      vy[u] = Y[e];
      vz[u] = Z[e];
      vm[u] = unit ? 0.f : M[e];
      vp[u] = ident ? (int32_t)e : P[e];  // the first level's order is the input's
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      if (u * 256 + (int)threadIdx.x >= blen) continue;
      const int64_t o = off + d[u];
      X2[o] = vx[u];
      Y2[o] = vy[u];
      Z2[o] = vz[u];
      if (!unit) M2[o] = vm[u];  // unit masses are not moved (never read again)
      P2[o] = vp[u];
    }
  }
}

// Children of non-degenerate splits form the next level
// (RCBForceTree.cxx:731-770: left = [offset, offset+is), right = rest).
__global__ void k_make_children(BNode *nodes, int64_t lb, int64_t le, int64_t nd,
                                int64_t next_base, Counters *ctr, Counters *ctr_next) {
  int64_t t = lb + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  // the next level's counters (the other of two, read by the host before
  // this level ends) start from zero
  if (blockIdx.x == 0 && threadIdx.x < sizeof(Counters) / sizeof(int64_t))
    reinterpret_cast<int64_t *>(ctr_next)[threadIdx.x] = 0;
  if (t >= le) return;
  BNode &n = nodes[t];
  if (n.kind != NODE_SPLIT) return;
  int64_t k = (int64_t)atomicAdd((unsigned long long *)&ctr->nsplit, 1ull);
  atomicAdd((unsigned long long *)&ctr->active_next,
            (unsigned long long)n.count);
  int64_t cl = next_base + 2 * k, cr = cl + 1;
  // which centre-of-mass paths the next level needs
  const int nbig = (n.is > CM_SMALL) + (n.count - n.is > CM_SMALL);
  if (nbig > 0) atomicAdd((unsigned long long *)&ctr->next_big, (unsigned long long)nbig);
  if (nbig < 2) atomicAdd((unsigned long long *)&ctr->next_small, (unsigned long long)(2 - nbig));
  const int nleaf = (n.is <= nd) + (n.count - n.is <= nd);
  if (nleaf > 0) atomicAdd((unsigned long long *)&ctr->next_leaf, (unsigned long long)nleaf);
  n.cl = (int32_t)cl;
  n.cr = (int32_t)cr;
  BNode &L = nodes[cl];
  BNode &R = nodes[cr];
  L.count = n.is;
  L.offset = n.offset;
  R.count = n.count - n.is;
  R.offset = n.offset + n.is;
  L.parent = R.parent = (int32_t)t;
  L.depth = R.depth = n.depth + 1;
  L.cl = L.cr = R.cl = R.cr = -1;
  L.nblk = R.nblk = 0;
}

// ------------------------------------------------------------ finishing
// Bottom-up: parent pseudo-particles from the children, left then right
// (RCBForceTree.cxx:856-889 with TDPTS = 1), allocating-node and leaf counts.
__global__ void k_bottom_up(BNode *nodes, int64_t lb, int64_t le, float ppc,
                            const int32_t *Pfin, const float *Min) {
  int64_t t = lb + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= le) return;
  BNode &n = nodes[t];
  if (n.kind == NODE_LEAF) {
    n.sa = 0;
    n.nleaf = 1;
    return;
  }
  n.tdr = ppc * pptdr_dev(n.xmin, n.xmax, n.xc);
  if (n.kind == NODE_DEGEN) {
    // both reserved children stay empty: no pseudo-particle mass
    n.ppm = 0.f;
    n.sa = 1;
    n.nleaf = 1;
    return;
  }
  const BNode &L = nodes[n.cl];
  const BNode &R = nodes[n.cr];
  float ppm = 0.f;
  const BNode *ch[2] = {&L, &R};
  for (int k = 0; k < 2; ++k) {
    const BNode &c = *ch[k];
    if (c.count > 0) {
      float cm = (c.count <= 1) ? Min[Pfin[c.offset]] : c.ppm;
      ppm += cm * 1.0f;
    }
  }
  n.ppm = ppm;
  n.sa = 1 + L.sa + R.sa;
  n.nleaf = L.nleaf + R.nleaf;
}

// Top-down: canonical ids (children get 1 + 2A and 2 + 2A, A = allocating
// nodes before the parent in preorder) and the right-first leaf rank.
__global__ void k_top_down(BNode *nodes, int64_t lb, int64_t le) {
  int64_t t = lb + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= le) return;
  BNode &n = nodes[t];
  if (n.kind != NODE_SPLIT) return;
  BNode &L = nodes[n.cl];
  BNode &R = nodes[n.cr];
  L.a = n.a + 1;
  R.a = n.a + 1 + L.sa;
  L.canon = 1 + 2 * n.a;
  R.canon = 2 + 2 * n.a;
  R.leaf_rank = n.leaf_rank;
  L.leaf_rank = n.leaf_rank + R.nleaf;
}

__global__ void k_write_canon(const BNode *nodes, int64_t nb,
                              const int64_t *box_cbase, const int32_t *box_of,
                              hsr_node *out) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nb) return;
  const BNode &n = nodes[t];
  int64_t base = box_cbase[box_of[t]];
  hsr_node o;
  o.count = n.count;
  o.offset = n.offset;
  o.cl = (n.kind == NODE_SPLIT) ? base + nodes[n.cl].canon : 0;
  o.cr = (n.kind == NODE_SPLIT) ? base + nodes[n.cr].canon : 0;
  o.ppm[0] = n.ppm;
  o.tdr = n.tdr;
  for (int d = 0; d < 3; ++d) {
    o.xmin[d] = n.xmin[d];
    o.xmax[d] = n.xmax[d];
    o.xc[d] = n.xc[d];
  }
  out[base + n.canon] = o;
  if (n.kind == NODE_DEGEN) {
    // the two reserved child slots (RCBForceTree.cxx:825-835): zeroed, with
    // the parent's bbox copied in
    hsr_node e;
    e.count = 0;
    e.offset = 0;
    e.cl = e.cr = 0;
    e.ppm[0] = 0.f;
    e.tdr = 0.f;
    for (int d = 0; d < 3; ++d) {
      e.xmin[d] = n.xmin[d];
      e.xmax[d] = n.xmax[d];
      e.xc[d] = 0.f;
    }
    out[base + 1 + 2 * n.a] = e;
    out[base + 2 + 2 * n.a] = e;
  }
}

__global__ void k_box_of(const BNode *nodes, int64_t nb, int32_t *box_of,
                         const int32_t *root_box) {
  // root_box[t] = box of root t or -1; walk parents
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nb) return;
  int64_t u = t;
  while (nodes[u].parent >= 0) u = nodes[u].parent;
  box_of[t] = root_box[u];
}

// ------------------------------------------------- quadrupole (TDPTS = 12)
// Hardin and Sloane's 12-point spherical 4-design (the icosahedron's
// vertices) as the reference tabulates it (RCBForceTree.cxx:229-272), in
// its order and rounded to float as its POSVEL_T arrays are.
#define A12 ((float)0.525731112119134)
#define B12 ((float)0.85065080835204)
__constant__ float kSph12X[12] = {0.f, 0.f, A12, -A12, B12, -B12, 0.f, 0.f, -A12, A12, -B12, B12};
__constant__ float kSph12Y[12] = {B12, B12, 0.f, 0.f, A12, A12, -B12, -B12, 0.f, 0.f, -A12, -A12};
__constant__ float kSph12Z[12] = {A12, -A12, B12, B12, 0.f, 0.f, -A12, A12, -B12, -B12, 0.f, 0.f};
#undef A12
#undef B12

// One term of pp<12> (RCBForceTree.cxx:549-565): the weight of a member at
// (x, y, z) with mass w for design point (pjx, pjy, pjz) of a node centred
// at xc with radius tdr.  odr2's product with the double literal 0.5 is one
// rounding of t2 * u / 2 in the reference; (t2 * 0.5f) * u rounds the same
// exact value once whenever t2 / 2 is a normal float, and when it is not,
// |odr2| < 2^-100 is absorbed by odr0 + odr1 ~ 1/12 either way.
__device__ __forceinline__ float pp12_term(float x, float y, float z, float w,
                                           const float *xc, float pjx, float pjy,
                                           float pjz, float tdr) {
  const float K = 12.0f;
  const float odr0 = 1 / K;
  const float xi = x - xc[0], yi = y - xc[1], zi = z - xc[2];
  const float ri = sqrtf(xi * xi + yi * yi + zi * zi);
  const float xj = pjx - xc[0], yj = pjy - xc[1], zj = pjz - xc[2];
  const float rj2 = xj * xj + yj * yj + zj * zj;
  float odr1 = 0.f, odr2 = 0.f;
  if (rj2 != 0.f) {
    const float rj = sqrtf(rj2);
    const float aij = (xi * xj + yi * yj + zi * zj) / (ri * rj);
    odr1 = (3 / K) * (ri / tdr) * aij;
    const float t2 = (5 / K) * (ri / tdr) * (ri / tdr);
    const float u = 3 * aij * aij - 1;
    odr2 = (t2 * 0.5f) * u;
  }
  return w * (odr0 + odr1 + odr2);
}

// pp<12> of every node of one BFS level (deepest level first), one thread
// per (node, design point): leaves with count > 12 from their members
// (RCBForceTree.cxx:787-797), split nodes from each child, left then right
// -- its members when count <= 12, else its 12 pseudo-particles
// (:856-889) -- accumulated in that order; degenerate nodes and leaves with
// count <= 12 keep 0.  Reads the canonical nodes and the particles in tree
// order; writes ppm12[12 * canonical id + j].
__global__ void k_pp12_level(const BNode *nodes, int64_t lb, int64_t le,
                             const int64_t *box_cbase, const int32_t *box_of,
                             hsr_node *cn, const float *xs, const float *ys,
                             const float *zs, const float *ms, float *ppm12) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t t = lb + g / 12;
  const int j = (int)(g % 12);
  if (t >= le) return;
  const BNode &b = nodes[t];
  const int64_t k = box_cbase[box_of[t]] + b.canon;
  const hsr_node &o = cn[k];
  const float pjx = o.tdr * kSph12X[j] + o.xc[0];
  const float pjy = o.tdr * kSph12Y[j] + o.xc[1];
  const float pjz = o.tdr * kSph12Z[j] + o.xc[2];
  float acc = 0.f;
  if (b.kind == NODE_LEAF) {
    if (o.count > 12)
      for (int64_t i = o.offset; i < o.offset + o.count; ++i)
        acc += pp12_term(xs[i], ys[i], zs[i], ms[i], o.xc, pjx, pjy, pjz, o.tdr);
  } else if (b.kind == NODE_SPLIT) {
    const int64_t ch[2] = {o.cl, o.cr};
    for (int q = 0; q < 2; ++q) {
      const hsr_node &c = cn[ch[q]];
      if (c.count <= 0) continue;
      if (c.count <= 12) {
        for (int64_t i = c.offset; i < c.offset + c.count; ++i)
          acc += pp12_term(xs[i], ys[i], zs[i], ms[i], o.xc, pjx, pjy, pjz, o.tdr);
      } else {
        const float *cm = ppm12 + 12 * ch[q];
        for (int i = 0; i < 12; ++i) {
          const float cx = c.tdr * kSph12X[i] + c.xc[0];
          const float cy = c.tdr * kSph12Y[i] + c.xc[1];
          const float cz = c.tdr * kSph12Z[i] + c.xc[2];
          acc += pp12_term(cx, cy, cz, cm[i], o.xc, pjx, pjy, pjz, o.tdr);
        }
      }
    }
  }
  ppm12[12 * k + j] = acc;
  if (j == 0) cn[k].ppm[0] = acc;  // the node table's mass slot holds ppm[0] of the 12
}

// Pseudo-particle records of TDPTS = 12: node k's design point j at
// tdr * design + xc with mass ppm12, at index count + 12 k + j.
__global__ void k_pseudo_pj12(int64_t n, int64_t nc, const hsr_node *cn, const float *ppm12,
                              PJ *pj) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= 12 * nc) return;
  const int64_t k = g / 12;
  const int j = (int)(g % 12);
  const hsr_node &o = cn[k];
  PJ r;
  r.x = o.tdr * kSph12X[j] + o.xc[0];
  r.y = o.tdr * kSph12Y[j] + o.xc[1];
  r.z = o.tdr * kSph12Z[j] + o.xc[2];
  r.m = ppm12[g];
  pj[n + g] = r;
}

__global__ void k_zero_f32(int64_t n, float *a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0.f;
}

// Pseudo-particle records (pppts with the TDPTS = 1 design point (0,0,0):
// position tdr*0 + xc, mass ppm), RCBForceTree.cxx:525-534, 1057-1060.
__global__ void k_pseudo_pj(int64_t n, int64_t nc, const hsr_node *cn,
                            PJ *pj) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nc) return;
  const hsr_node &o = cn[k];
  PJ r;
  r.x = o.tdr * 0.0f + o.xc[0];
  r.y = o.tdr * 0.0f + o.xc[1];
  r.z = o.tdr * 0.0f + o.xc[2];
  r.m = o.ppm[0];
  pj[n + k] = r;
}

__global__ void k_init_roots(BNode *nodes, int64_t nboxes,
                             const int64_t *box_off, const float *min_loc,
                             const float *max_loc, int32_t *P, int64_t n) {
  int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b < nboxes) {
    BNode &r = nodes[b];
    r.count = box_off[b + 1] - box_off[b];
    r.offset = box_off[b];
    r.parent = -1;
    r.cl = r.cr = -1;
    r.depth = 0;
    r.kind = NODE_LEAF;
    r.nblk = 0;
    r.canon = 0;
    r.a = 0;
    r.leaf_rank = 0;
    r.ppm = 0.f;
    r.tdr = 0.f;
    for (int d = 0; d < 3; ++d) {
      r.xmin[d] = min_loc[3 * b + d];
      r.xmax[d] = max_loc[3 * b + d];
      r.xc[d] = 0.f;
    }
  }
}

// flag[0] = 0 unless every mass is exactly 1.0f (flag[0] preset to 1 by
// k_set_flag); every writer stores the same value.
__global__ void k_set_flag(int32_t *flag) { flag[0] = 1; }

__global__ void __launch_bounds__(256)
k_unit_check(int64_t n, const float *m, int32_t *flag) {
  bool ok = true;
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  // float4 loads from the first 16-byte boundary, scalar head and tail
  const int64_t head = std::min<int64_t>(n, (int64_t)(((16 - ((uintptr_t)m & 15)) & 15) / 4));
  const int64_t nv = (n - head) / 4;
  const float4 *m4 = reinterpret_cast<const float4 *>(m + head);
  for (int64_t i = g; i < nv; i += stride) {
    const float4 v = m4[i];
    ok = ok && v.x == 1.0f && v.y == 1.0f && v.z == 1.0f && v.w == 1.0f;
  }
  if (g < head) ok = ok && m[g] == 1.0f;
  for (int64_t i = head + 4 * nv + g; i < n; i += stride) ok = ok && m[i] == 1.0f;
  if (!__all(ok) && lane_id() == 0) flag[0] = 0;
}

__global__ void k_zero_counters(Counters *ctr) {
  int64_t *p = reinterpret_cast<int64_t *>(ctr);
  if (threadIdx.x < sizeof(Counters) / sizeof(int64_t)) p[threadIdx.x] = 0;
}

inline unsigned nblk_for(int64_t n, int t) {
  return (unsigned)((n + t - 1) / t);
}

}  // namespace

hsr_status tree_build_device(hsr_ctx *c, int64_t nboxes,
                             const int64_t *box_off_h, const float *min_loc_h,
                             const float *max_loc_h, const float *x,
                             const float *y, const float *z, const float *m) {
  hipStream_t st = c->stream;
  TreeState &T = c->tree;
  Scratch &S = c->s;
  int64_t n = box_off_h[nboxes];
  T.n = n;
  T.nboxes = nboxes;
  const int64_t nd = c->p.nd;

  // node table capacity: grows between levels when needed
  int64_t cap = std::max<int64_t>(
      64 + 2 * nboxes, (((int64_t)((double)n / (double)nd)) << std::max<int64_t>(c->p.ds, 1)) + 1 + nboxes);
  HSR_HIP(dbuf_reserve(T.bnodes, sizeof(BNode) * cap));
  size_t nn = (size_t)std::max<int64_t>(n, 1);
  for (DBuf *b : {&S.x0, &S.y0, &S.z0, &S.m0, &S.x1, &S.y1, &S.z1, &S.m1})
    HSR_HIP(dbuf_reserve(*b, sizeof(float) * nn));
  HSR_HIP(dbuf_reserve(S.p0, sizeof(int32_t) * nn));
  HSR_HIP(dbuf_reserve(S.p1, sizeof(int32_t) * nn));
  HSR_HIP(dbuf_reserve(T.pfin, sizeof(int32_t) * nn));
  HSR_HIP(dbuf_reserve(S.a, sizeof(int32_t) * nn));  // idx
  HSR_HIP(dbuf_reserve(S.b, sizeof(int32_t) * nn));  // dest
  HSR_HIP(dbuf_reserve(S.c, sizeof(int32_t) * nn));  // jmp
  HSR_HIP(dbuf_reserve(S.rights, sizeof(int32_t) * nn));
  HSR_HIP(dbuf_reserve(S.counters, 2 * sizeof(Counters) + 64));
  HSR_HIP(dbuf_reserve(S.d, sizeof(int64_t) * (nboxes + 1) + sizeof(float) * 6 * nboxes + 64));
  // (the 64 spare bytes after the box table hold the unit-mass flag)

  float *X[2] = {dptr<float>(S.x0), dptr<float>(S.x1)};
  float *Y[2] = {dptr<float>(S.y0), dptr<float>(S.y1)};
  float *Z[2] = {dptr<float>(S.z0), dptr<float>(S.z1)};
  float *Mm[2] = {dptr<float>(S.m0), dptr<float>(S.m1)};
  int32_t *P[2] = {dptr<int32_t>(S.p0), dptr<int32_t>(S.p1)};
  int32_t *Pfin = dptr<int32_t>(T.pfin);
  int32_t *idx = dptr<int32_t>(S.a);
  int32_t *dest = dptr<int32_t>(S.b);
  int32_t *jmp = dptr<int32_t>(S.c);
  int32_t *rights = dptr<int32_t>(S.rights);
  Counters *ctr2 = dptr<Counters>(S.counters);  // two sets, alternating by level
  // sorted arrays, written as nodes stop splitting
  for (DBuf *bb : {&T.xs, &T.ys, &T.zs, &T.ms})
    HSR_HIP(dbuf_reserve(*bb, sizeof(float) * nn));
  float *xs_ = dptr<float>(T.xs), *ys_ = dptr<float>(T.ys), *zs_ = dptr<float>(T.zs),
        *ms_ = dptr<float>(T.ms);

  // box table
  int64_t *d_box_off = dptr<int64_t>(S.d);
  float *d_min = reinterpret_cast<float *>(d_box_off + nboxes + 1);
  float *d_max = d_min + 3 * nboxes;
  // Unit masses (map2 sets every mass to 1 before the tree build): the mass
  // sums are the counts, the addends w x are x, and the partition need not
  // move the masses.  Every result is the same as with the general path.
  int32_t *d_unit = reinterpret_cast<int32_t *>(d_max + 3 * nboxes);
  k_set_flag<<<1, 1, 0, st>>>(d_unit);
  if (n > 0)
    k_unit_check<<<(unsigned)std::min<int64_t>(nblk_for(n, 256), 4096), 256, 0, st>>>(n, m, d_unit);
  int32_t unit_h = 0;
  HSR_HIP(hipMemcpyAsync(&unit_h, d_unit, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(stream_wait(st));
  const bool unit = unit_h != 0;
  T.unit = unit;
  const bool unit_ppm = unit;  // leaf pseudo masses set by classify_node (k_part_plan)
  HSR_HIP(hipMemcpyAsync(d_box_off, box_off_h, sizeof(int64_t) * (nboxes + 1), hipMemcpyHostToDevice, st));
  HSR_HIP(hipMemcpyAsync(d_min, min_loc_h, sizeof(float) * 3 * nboxes, hipMemcpyHostToDevice, st));
  HSR_HIP(hipMemcpyAsync(d_max, max_loc_h, sizeof(float) * 3 * nboxes, hipMemcpyHostToDevice, st));
  BNode *nodes = dptr<BNode>(T.bnodes);
  k_init_roots<<<nblk_for(nboxes, 64), 64, 0, st>>>(nodes, nboxes, d_box_off, d_min, d_max, P[0], n);
  const unsigned nq = unit ? 3u : 4u;  // sums formed from addends

  T.level_start.assign(1, 0);
  T.level_start.push_back(nboxes);
  int cur = 0;
  int64_t active = n;
  Counters hc;
  int depth = 0;
  bool run_big = true, run_small = true, run_leaf = true;
  for (;;) {
    int64_t lb = T.level_start[depth], le = T.level_start[depth + 1];
    int64_t nl = le - lb;
    if (nl == 0) break;
    // capacity for the children of this level (at most 2 per node)
    if ((le + 2 * nl) > cap) {
      int64_t ncap = std::max<int64_t>(cap * 2, le + 2 * nl + 64);
      DBuf nb;
      HSR_HIP(dbuf_reserve(nb, sizeof(BNode) * ncap));
      HSR_HIP(hipMemcpyAsync(nb.p, T.bnodes.p, sizeof(BNode) * le, hipMemcpyDeviceToDevice, st));
      HSR_HIP(stream_wait(st));
      dbuf_free(T.bnodes);
      T.bnodes = nb;
      cap = ncap;
      nodes = dptr<BNode>(T.bnodes);
    }
    Counters *ctr = ctr2 + (depth & 1);
    if (depth == 0) k_zero_counters<<<1, 64, 0, st>>>(ctr);
    // the first level reads the caller's arrays; its partition writes the
    // first copy (buffer 1), and the levels after it alternate
    const float *Xr = depth == 0 ? x : X[cur], *Yr = depth == 0 ? y : Y[cur];
    const float *Zr = depth == 0 ? z : Z[cur], *Mr = depth == 0 ? m : Mm[cur];
    // ---- centre of mass
    int64_t cm_cap = active / CM_BLK + active / CM_SMALL + 2;
    HSR_HIP(dbuf_reserve(S.blocks, sizeof(CmBlock) * cm_cap));
    CmBlock *cmb = dptr<CmBlock>(S.blocks);
    HSR_HIP(dbuf_reserve(S.cms, sizeof(sd_t) * 4 * (size_t)nl));
    sd_t *cms = dptr<sd_t>(S.cms);
    // Kernels over the big nodes (count > CM_SMALL) and their blocks loop
    // over a grid of at most gbig workgroups: deep levels hold many nodes
    // but few big ones, and empty workgroups would cost their dispatch.
    const unsigned gbig = (unsigned)std::max(1, c->num_cus) * 32;
    const unsigned gnl = (unsigned)std::min<int64_t>(nl, gbig);
    HSR_HIP(dbuf_reserve(S.cmfail, sizeof(int32_t) * 4 * (size_t)nl));
    int32_t *cm_list = dptr<int32_t>(S.cmfail);
    // the small-node and big-node paths are launched only for levels that
    // hold such nodes (the previous level's children counts; both at the
    // first level); each path's kernels skip the other's nodes
    if (run_small) {
      k_cm_small<<<dim3((unsigned)nl, nq), 64, 0, st>>>(nodes, lb, le, unit, Xr, Yr, Zr, Mr,
                                                       cms, cm_list, ctr);
      k_cm_small_ordered<<<std::min<unsigned>((unsigned)std::max<int64_t>(1, nl), 4 * gbig), 64, 0, st>>>(
          nodes, lb, unit, Xr, Yr, Zr, Mr, cms, cm_list, ctr);
    }
    HSR_HIP(dbuf_reserve(S.cnode, sizeof(int32_t) * cm_cap));
    int32_t *cnode = dptr<int32_t>(S.cnode);
    HSR_HIP(dbuf_reserve(S.cmx, sizeof(int32_t) * (size_t)nl));
    int32_t *cmflags = dptr<int32_t>(S.cmx);
    if (run_big) {
      k_cm_plan<<<nblk_for(nl, 64), 64, 0, st>>>(nodes, lb, le, cnode, ctr);
      unsigned cmg = (unsigned)std::max<int64_t>(1, active / CM_BLK + active / CM_SMALL + 1);
      cmg = std::min(cmg, gbig);
      if (unit)
        k_cm_blk_stats<true><<<cmg, 64, 0, st>>>(nodes, cnode, cmb, ctr, Xr, Yr, Zr, Mr);
      else
        k_cm_blk_stats<false><<<cmg, 64, 0, st>>>(nodes, cnode, cmb, ctr, Xr, Yr, Zr, Mr);
      const unsigned cmr = std::max(1u, std::min<unsigned>(gbig, (unsigned)(cm_cap / 64 + 1)));
      k_cm_chunk<<<cmr, 64, 0, st>>>(nodes, cnode, cmb, ctr);
      k_cm_node<<<gnl, 64, 0, st>>>(nodes, lb, le, cmb, cmflags, cms);
      k_cm_blk_maps<<<dim3(cmg, nq), 64, 0, st>>>(nodes, cnode, cmb, ctr, unit, Xr, Yr, Zr, Mr, lb,
                                                  cmflags);
      k_cm_win_maps<<<cmg, 64, 0, st>>>(nodes, cnode, cmb, ctr);
      k_cm_blk_chain<<<dim3(gnl, nq), 64, 0, st>>>(nodes, lb, le, unit, cmb, Xr, Yr, Zr, Mr, ctr, cms,
                                                   cmflags);
    }
    // ---- centre of mass, class and partition blocks (a lane per node),
    // finalize plain leaves
    int64_t pcap = active / PB + nl + 2;
    HSR_HIP(dbuf_reserve(S.pblocks, sizeof(PartBlock) * pcap));
    PartBlock *pb = dptr<PartBlock>(S.pblocks);
    HSR_HIP(dbuf_reserve(S.pnode, sizeof(int32_t) * pcap));
    int32_t *pnode = dptr<int32_t>(S.pnode);
    k_part_plan<<<nblk_for(nl, 64), 64, 0, st>>>(nodes, lb, le, unit, cms, nd, c->p.ppc, unit_ppm, pnode, ctr);
    if (!unit_ppm && run_leaf) k_leaf_ppm<<<gnl, 64, 0, st>>>(nodes, lb, le, nd, Mr);
    if (run_leaf)
      k_finalize_leaves<<<(unsigned)nl, 64, 0, st>>>(nodes, lb, le, NODE_LEAF, depth == 0, P[cur], Pfin, unit,
                                                     Xr, Yr, Zr, Mr, xs_, ys_, zs_, ms_);
    // ---- partition of splitting nodes
    const unsigned pg = (unsigned)std::max<int64_t>(1, pcap);
    HSR_HIP(dbuf_reserve(S.pmask, sizeof(uint64_t) * (PB / 64) * (size_t)pcap));
    uint64_t *pmask = dptr<uint64_t>(S.pmask);
    k_part_count<<<pg, 256, 0, st>>>(nodes, pnode, pb, ctr, Xr, Yr, Zr, pmask);
    if (active / PB > 64 * nl)  // over 64 blocks per node on average
      k_part_scan<PFX_WAVES><<<std::min<unsigned>((unsigned)nl, 4u * (unsigned)std::max(1, c->num_cus)), PFX_T,
                               0, st>>>(nodes, lb, le, pb, ctr);
    else
      k_part_scan<1><<<std::min<unsigned>((unsigned)nl, gbig), 64, 0, st>>>(nodes, lb, le, pb, ctr);
    // children depend only on the partition counts: create them now so a
    // single host synchronisation per level reads every counter
    k_make_children<<<nblk_for(nl, 64), 64, 0, st>>>(nodes, lb, le, nd, le, ctr, ctr2 + ((depth + 1) & 1));
    k_part_rank<<<pg, 256, 0, st>>>(nodes, pnode, pb, ctr, pmask, idx, rights, dest);
    k_part_chain<<<pg, 256, 0, st>>>(nodes, pnode, ctr, idx, rights, dest, ctr,
                                     c->tuning.chain_hops > 0 ? c->tuning.chain_hops : MAXHOP);
    HSR_HIP(hipMemcpyAsync(&hc, ctr, sizeof(Counters), hipMemcpyDeviceToHost, st));
    HSR_HIP(stream_wait(st));
    if (hc.unresolved > 0) {
      k_part_jump_init<<<pg, 256, 0, st>>>(nodes, pnode, ctr, idx, jmp);
      for (int r = 0; r < 32; ++r)
        k_part_jump_round<<<pg, 256, 0, st>>>(nodes, pnode, ctr, jmp);
      k_part_jump_final<<<pg, 256, 0, st>>>(nodes, pnode, ctr, jmp, dest);
    }
    k_part_scatter<<<pg, 256, 0, st>>>(nodes, pnode, ctr, unit, depth == 0, dest, Xr, Yr, Zr, Mr, P[cur],
                                        X[cur ^ 1], Y[cur ^ 1], Z[cur ^ 1], Mm[cur ^ 1], P[cur ^ 1]);
    if (hc.ndegen > 0)
      k_finalize_leaves<<<(unsigned)nl, 64, 0, st>>>(nodes, lb, le, NODE_DEGEN, false, P[cur ^ 1], Pfin, unit,
                                                     X[cur ^ 1], Y[cur ^ 1], Z[cur ^ 1], Mm[cur ^ 1], xs_,
                                                     ys_, zs_, ms_);
    HSR_HIP(hipGetLastError());
#ifdef HSR_CM_STATS
    fprintf(stderr, "[cm] level %d nodes %lld blocks %lld scans %lld fallback %lld small ordered %lld\n",
            depth, (long long)nl, (long long)hc.cm_blocks, (long long)hc.pad[0], (long long)hc.pad[1],
            (long long)hc.cm_fail);
#endif
    T.level_start.push_back(le + 2 * hc.nsplit);
    active = hc.active_next;
    run_big = hc.next_big > 0;
    run_small = hc.next_small > 0;
    run_leaf = hc.next_leaf > 0;
    cur ^= 1;
    ++depth;
  }
  int nlev = (int)T.level_start.size() - 1;  // last level is empty
  T.nnodes_bfs = T.level_start.back();
  T.max_depth = std::max(0, nlev - 2);  // deepest non-empty level index
  int64_t nb = T.nnodes_bfs;

  // ---- bottom-up, top-down
  const float *Min = m;
  for (int l = nlev - 1; l >= 0; --l) {
    int64_t lb = T.level_start[l], le = T.level_start[l + 1];
    if (le > lb)
      k_bottom_up<<<nblk_for(le - lb, 64), 64, 0, st>>>(nodes, lb, le, c->p.ppc, Pfin, Min);
  }
  for (int l = 0; l < nlev; ++l) {
    int64_t lb = T.level_start[l], le = T.level_start[l + 1];
    if (le > lb) k_top_down<<<nblk_for(le - lb, 64), 64, 0, st>>>(nodes, lb, le);
  }
  // per-box canonical bases: box b holds 1 + 2 * sa(root_b) nodes
  std::vector<BNode> roots(nboxes);
  HSR_HIP(hipMemcpyAsync(roots.data(), nodes, sizeof(BNode) * nboxes, hipMemcpyDeviceToHost, st));
  HSR_HIP(stream_wait(st));
  std::vector<int64_t> cbase(nboxes + 1, 0);
  for (int64_t b = 0; b < nboxes; ++b) cbase[b + 1] = cbase[b] + 1 + 2 * roots[b].sa;
  T.nnodes = cbase[nboxes];
  T.box_root.assign(cbase.begin(), cbase.end());
  HSR_HIP(dbuf_reserve(S.e, sizeof(int64_t) * (nboxes + 1) + sizeof(int32_t) * (nb + nboxes) + 64));
  int64_t *d_cbase = dptr<int64_t>(S.e);
  int32_t *root_box = reinterpret_cast<int32_t *>(d_cbase + nboxes + 1);
  HSR_HIP(hipMemcpyAsync(d_cbase, cbase.data(), sizeof(int64_t) * (nboxes + 1), hipMemcpyHostToDevice, st));
  {
    std::vector<int32_t> rb(nboxes);
    for (int64_t b = 0; b < nboxes; ++b) rb[b] = (int32_t)b;
    HSR_HIP(hipMemcpyAsync(root_box, rb.data(), sizeof(int32_t) * nboxes, hipMemcpyHostToDevice, st));
    HSR_HIP(stream_wait(st));
  }
  HSR_HIP(dbuf_reserve(S.f, sizeof(int32_t) * (nb + 1)));
  int32_t *box_of = dptr<int32_t>(S.f);
  k_box_of<<<nblk_for(nb, 256), 256, 0, st>>>(nodes, nb, box_of, root_box);
  HSR_HIP(dbuf_reserve(T.canon_nodes, sizeof(hsr_node) * std::max<int64_t>(T.nnodes, 1)));
  hsr_node *cn = dptr<hsr_node>(T.canon_nodes);
  k_write_canon<<<nblk_for(nb, 256), 256, 0, st>>>(nodes, nb, d_cbase, box_of, cn);

  // ---- j-particle records of the particles (the sorted arrays are done)
  const int tdpts = c->p.tdpts;
  HSR_HIP(dbuf_reserve(T.pj, sizeof(PJ) * (n + tdpts * T.nnodes + 1)));
  if (n > 0)
    k_pj_particles<<<nblk_for(n, 256), 256, 0, st>>>(n, xs_, ys_, zs_, ms_, dptr<PJ>(T.pj));
  if (tdpts == 12) {
    // quadrupole pseudo-particles, deepest BFS level first; the orphan
    // slots of degenerate nodes keep the zeros
    HSR_HIP(dbuf_reserve(T.ppm12, sizeof(float) * 12 * std::max<int64_t>(T.nnodes, 1)));
    float *ppm12 = dptr<float>(T.ppm12);
    if (T.nnodes > 0) k_zero_f32<<<nblk_for(12 * T.nnodes, 256), 256, 0, st>>>(12 * T.nnodes, ppm12);
    for (int l = nlev - 1; l >= 0; --l) {
      int64_t lb = T.level_start[l], le = T.level_start[l + 1];
      if (le > lb)
        k_pp12_level<<<nblk_for(12 * (le - lb), 256), 256, 0, st>>>(
            nodes, lb, le, d_cbase, box_of, cn, dptr<float>(T.xs), dptr<float>(T.ys), dptr<float>(T.zs),
            dptr<float>(T.ms), ppm12);
    }
    if (T.nnodes > 0)
      k_pseudo_pj12<<<nblk_for(12 * T.nnodes, 256), 256, 0, st>>>(n, T.nnodes, cn, ppm12, dptr<PJ>(T.pj));
  } else if (T.nnodes > 0) {
    k_pseudo_pj<<<nblk_for(T.nnodes, 256), 256, 0, st>>>(n, T.nnodes, cn, dptr<PJ>(T.pj));
  }
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}
