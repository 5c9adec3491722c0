// force.hip -- short-range leaf kernel for gfx950 (FP32 VALU; no MFMA: the
// pair force is not a dense contraction).
//
// Pair law (BGQStep16.c:167-187, the 5th-order polynomial of ForceLaw.cxx
// :107-116 at POLY_ORDER 5):
//     f(r2) = m_j * [ (r2 + rsm^2)^(-3/2) - P(r2) ] * [r2 < rc2] (* [r2 > 0])
//     F_i  += f * (x_j - x_i)
// summed over every entry j of leaf i's interaction list (nbody1,
// RCBForceTree.cxx:599-619).  The polynomial is evaluated in s = r2 + rsm^2
// (coefficients re-expanded on the host), so s serves the power, the
// polynomial and the cutoff test.
//
// Work: one persistent launch walks a queue of work items -- a listed leaf,
// or a chunk of a long list -- sorted by decreasing pair count.  A wave
// takes an item: the leaf's i-particles sit in registers, R per lane, in G
// lane groups; the j-entries are staged in LDS per tile, each group owning a
// contiguous, padded slice (stride = 1 mod 8 entries, so the groups' 16-byte
// reads fall in distinct banks).  Chunk partials of long lists are summed in
// chunk order afterwards, so results are bitwise reproducible.
#include "hsr_internal.h"
#include "wave.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace {

// Kernel shapes: RR i-particles per lane, UU j entries per unrolled step,
// TT j entries staged per LDS tile.  The default (2, 2, 256) is the fastest
// measured at 256^3; the others are built for the tuner (hsr_tuning
// i_per_lane, unroll, tile).
constexpr int DEF_R = 2, DEF_U = 2, DEF_T = 256;
constexpr int GMAX = 16;      // at most 16 lane groups (>= 4 lanes each)
// Passes hold a power of two of at least 4 R targets, or fewer than 4 R, so
// the group count G = 64 / max(ceil(targets / R), 4) is 1, 2, 4, 8 or 16;
// the staged slices of a full tile must fit for each of them.  A group's
// slice holds its unmasked entries (a multiple of UU) and its masked ones
// rounded up to UU: at most ceil(entries / G) + UU - 1 slots.
template <int UU>
__host__ __device__ constexpr int slice_slots(int nt, int G) { return (nt + G - 1) / G + UU; }
template <int UU, int TT>
constexpr int staged_entries(int G) {
  const int tg = slice_slots<UU>(TT, G);
  return G * (tg + ((9 - (tg & 7)) & 7));
}
template <int UU, int TT>
constexpr int lds_entries() {
  int m = 0;
  for (int G = 1; G <= GMAX; G *= 2) m = staged_entries<UU, TT>(G) > m ? staged_entries<UU, TT>(G) : m;
  return m;
}

struct Item {
  int32_t leaf;     // listed-leaf index
  int32_t chunk;    // chunk index within the leaf's list
  int64_t e0, e1;   // list entry range
  int64_t pbase;    // partial-force offset (-1: write F directly)
};

struct LawArgs {
  float c[7];      // polynomial in s (degree 5 or 6)
  float rcs;       // s cutoff: pairs with s < rcs contribute
  float eps2;
  int guard;
  int cull;        // drop entries farther than the cutoff from the leaf's box
  float cull2;     // squared box distance beyond which every pair is masked
  int pmax;        // targets per pass (64 R or 32 R)
  float in2;       // squared farthest box distance below which every pair of
                   // the pass is inside the cutoff (the unmasked class)
  int unmasked;    // 1: unit-mass entries inside the cutoff skip mask and mass
  // ForceLawSR with FGridEvalFit or FGridEvalInterp: f = s^-3/2 - g(r2),
  // pairs with r2 < rc2 contribute
  float rc2;
  float b, cc, d, e, ff, g, h, l, rmax, f0;  // FGrid constants (ForceLaw.cxx:23-33)
  const float4 *bins;                        // {r2[i], f[i], f[i+1] - f[i], 0}
  float t_r2min, t_r2max, t_oodr2;
  int t_last;                                // last bin index (n_interp - 2)
};

// FGrid::fgor(r), the fitted grid force over r (ForceLaw.cxx:39-51), in
// float with the device's coshf, tanhf and expf (the reference mixes a few
// double products in; the difference is a rounding of fgor, far inside the
// force tolerance).  Entries padded far away keep every term finite.
__device__ __forceinline__ float fgor_fit(const LawArgs &L, float r) {
  const float r2 = r * r, r4 = r2 * r2, r6 = r4 * r2;
  const float br = L.b * r;
  const float ch = coshf(br);
  const float t = tanhf(br) - br / ch / ch +
                  L.cc * r * r2 * (1.f + L.d * r2) * expf(-L.d * r2) +
                  L.e * r2 * (L.ff * r2 + L.g * r4 + L.l * r6) * expf(-L.h * r2);
  const float rdiv = r + (r <= 0.f ? 1.f : 0.f);
  const float v = t / (rdiv * rdiv * rdiv) + (r <= 0.f ? L.f0 : 0.f);
  return (r <= L.rmax && r >= 0.f) ? v : 0.f;
}

// FGridEvalInterp::eval(r2) (ForceLaw.cxx:160-171): linear interpolation in
// the caller's table, zero outside (r2min, r2max).  The bin index is kept
// below the last point, which the reference's index can only reach through
// rounding just under r2max.
__device__ __forceinline__ float fgor_interp(const LawArgs &L, float r2) {
  const bool in = r2 > L.t_r2min && r2 < L.t_r2max;
  int k = in ? (int)((r2 - L.t_r2min) * L.t_oodr2) : 0;
  k = min(k, L.t_last);
  const float4 q = L.bins[k];
  const float v = q.y + (r2 - q.x) * L.t_oodr2 * q.z;
  return in ? v : 0.f;
}

// Squared distance from p to the box [lo, hi] (0 inside).
__device__ __forceinline__ float box_d2(const float *lo, const float *hi, float x, float y,
                                        float z) {
  const float dx = fmaxf(0.f, fmaxf(lo[0] - x, x - hi[0]));
  const float dy = fmaxf(0.f, fmaxf(lo[1] - y, y - hi[1]));
  const float dz = fmaxf(0.f, fmaxf(lo[2] - z, z - hi[2]));
  return dx * dx + dy * dy + dz * dz;
}

// Squared distance from p to the farthest point of the box [lo, hi].
__device__ __forceinline__ float box_far2(const float *lo, const float *hi, float x, float y,
                                          float z) {
  const float dx = fmaxf(fabsf(x - lo[0]), fabsf(x - hi[0]));
  const float dy = fmaxf(fabsf(y - lo[1]), fabsf(y - hi[1]));
  const float dz = fmaxf(fabsf(z - lo[2]), fabsf(z - hi[2]));
  return dx * dx + dy * dy + dz * dz;
}

template <int ORDER>
__device__ __forceinline__ float poly_tail(const LawArgs &L, float s) {
  // q(s) = c1 + c2 s + ... ; f = y^3 - c0 - s q(s)
  float q;
  if (ORDER == 6) {
    q = fmaf(L.c[6], s, L.c[5]);
    q = fmaf(q, s, L.c[4]);
  } else {
    q = fmaf(L.c[5], s, L.c[4]);
  }
  q = fmaf(q, s, L.c[3]);
  q = fmaf(q, s, L.c[2]);
  q = fmaf(q, s, L.c[1]);
  return q;
}

// Targets of the next pass over a leaf: 64 R while that many remain, then
// the largest power of two of at least 4 R that fits, then the rest.
template <int R>
__device__ __forceinline__ int64_t pass_size(int64_t rem, int pmax) {
  if (rem >= pmax) return pmax;
  for (int64_t c = pmax >> 1; c >= 4 * R; c >>= 1)
    if (rem >= c) return c;
  return rem;
}

// Waves per SIMD the register allocation of the leaf kernel targets (0: the
// compiler's choice: 94 VGPRs, 5 waves, with the four pairs of an unrolled
// step interleaved; a cap of 6 waves fits 80 VGPRs but serialises them and
// measured 2.7% slower at 256^3).
#ifndef HSR_FORCE_WAVES
#define HSR_FORCE_WAVES 0
#endif
#if HSR_FORCE_WAVES > 0
#define HSR_FORCE_ATTR __attribute__((amdgpu_waves_per_eu(HSR_FORCE_WAVES, HSR_FORCE_WAVES)))
#else
#define HSR_FORCE_ATTR
#endif

// One pair: F_i += m_j f(r2) (x_j - x_i).  MASKED applies the cutoff (and
// the r2 > 0 guard) and the source mass; unmasked entries are known to have
// unit mass and every pair inside the cutoff, where both are exact no-ops.
template <int LAWK, int ORDER, bool GUARD, bool MASKED>
__device__ __forceinline__ void pair(const LawArgs &L, float c0, float rcs, const float4 p,
                                     float xi, float yi, float zi, float &ax, float &ay,
                                     float &az) {
  const float dx = p.x - xi;
  const float dy = p.y - yi;
  const float dz = p.z - zi;
  float f;
  if constexpr (LAWK == LAWK_POLY) {
    float s = fmaf(dx, dx, L.eps2);
    s = fmaf(dy, dy, s);
    s = fmaf(dz, dz, s);
    const float y = __builtin_amdgcn_rsqf(s);
    const float y2 = y * y;
    const float t3 = fmaf(y2, y, -c0);
    const float q = poly_tail<ORDER>(L, s);
    f = fmaf(-s, q, t3);
    if constexpr (MASKED) {
      float mm = (s < rcs) ? p.w : 0.f;
      if (GUARD) mm = (s > L.eps2) ? mm : 0.f;
      f *= mm;
    }
  } else {
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float s = r2 + L.eps2;
    const float y = __builtin_amdgcn_rsqf(s);
    const float gr = LAWK == LAWK_FIT ? fgor_fit(L, __builtin_sqrtf(r2)) : fgor_interp(L, r2);
    f = y * y * y - gr;
    if constexpr (MASKED) {
      float mm = (r2 < L.rc2) ? p.w : 0.f;
      if (GUARD) mm = (r2 > 0.f) ? mm : 0.f;
      f *= mm;
    }
  }
  ax = fmaf(f, dx, ax);
  ay = fmaf(f, dy, ay);
  az = fmaf(f, dz, az);
}

template <int LAWK, int ORDER, bool GUARD, int R, int UNROLL, int TILE>
__global__ void __launch_bounds__(64) HSR_FORCE_ATTR
k_force(const hsr_node *__restrict__ cn, const int32_t *__restrict__ leaves,
        const Item *__restrict__ items, const int64_t *__restrict__ nitems_p,
        const int32_t *__restrict__ jidx, const PJ *__restrict__ pj,
        const int32_t *__restrict__ iperm, LawArgs L,
        unsigned long long *__restrict__ qctr, float *__restrict__ fx,
        float *__restrict__ fy, float *__restrict__ fz, float *__restrict__ part,
        unsigned long long *__restrict__ evals) {
  static_assert(R == 1 || R == 2 || R == 4, "i-particles per lane");
  static_assert(UNROLL >= 1 && UNROLL <= 8, "j entries per unrolled step");
  static_assert(TILE % 64 == 0 && TILE <= 512, "tile of 64-entry chunks");
  constexpr int NCH = TILE / 64;  // 64-entry chunks per tile
  constexpr int LDS_ENT = lds_entries<UNROLL, TILE>();
  // the tile doubles as the reduction buffer once the j-walk is done
  __shared__ float4 tile[LDS_ENT];
  static_assert(3 * 64 * R <= 4 * LDS_ENT, "reduction buffer fits the tile");
  float (*red)[64 * R] = reinterpret_cast<float (*)[64 * R]>(&tile[0]);
  const int l = lane_id();
  const int64_t nitems = *nitems_p;
  const float c0 = L.c[0], rcs = L.rcs;
  for (;;) {
    int64_t it = 0;
    if (l == 0) it = (int64_t)atomicAdd(qctr, 1ull);
    it = (int64_t)readlane_u64((uint64_t)it, 0);
    if (it >= nitems) break;
    const Item I = items[it];
    const hsr_node &nd = cn[leaves[I.leaf]];
    const int64_t off = nd.offset, cnt = nd.count;
    // padding entries: far enough to be masked, near enough to keep every
    // intermediate finite
    const float4 pad = make_float4(nd.xc[0] + 1000.f, nd.xc[1], nd.xc[2], 0.f);
    int64_t ev = 0;   // pairs evaluated for this item (culled entries excluded)
    int64_t evu = 0;  // of which unmasked
    // passes over the targets in power-of-two chunks (128, 64, ..., 8, rest)
    // so that every pass fills its groups of lpg = chunk / R lanes
    for (int64_t ib = 0, nib = 0; ib < cnt; ib += nib) {
      nib = pass_size<R>(cnt - ib, L.pmax);
      const int ni = (int)nib;
      const int lpg = max((ni + R - 1) / R, 64 / GMAX);
      const int G = max(1, 64 / lpg);
      const int g = l / lpg, tq = l - g * lpg;
      const bool active = g < G;
      float xi[R], yi[R], zi[R], ax[R], ay[R], az[R];
      // the pass's targets (in the leaf's target order) and their box
      float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int ii = tq + r * lpg;
        const bool v = active && ii < ni;
        const PJ p = pj[iperm[off + ib + (v ? ii : 0)]];
        xi[r] = v ? p.x : 1e30f;
        yi[r] = v ? p.y : 1e30f;
        zi[r] = v ? p.z : 1e30f;
        ax[r] = ay[r] = az[r] = 0.f;
        if (v) {
          bmin[0] = fminf(bmin[0], p.x); bmax[0] = fmaxf(bmax[0], p.x);
          bmin[1] = fminf(bmin[1], p.y); bmax[1] = fmaxf(bmax[1], p.y);
          bmin[2] = fminf(bmin[2], p.z); bmax[2] = fmaxf(bmax[2], p.z);
        }
      }
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          bmin[d] = fminf(bmin[d], __shfl_xor(bmin[d], o));
          bmax[d] = fmaxf(bmax[d], __shfl_xor(bmax[d], o));
        }
      }
      // the box is uniform across the wave: keep it in scalar registers
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        bmin[d] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, bmin[d])));
        bmax[d] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, bmax[d])));
      }
      const int lg = __builtin_ctz(G);  // G is a power of two
      for (int64_t t0 = I.e0; t0 < I.e1; t0 += TILE) {
        const int nt = (int)min<int64_t>(TILE, I.e1 - t0);
        int tgmax = slice_slots<UNROLL>(nt, G);
        const int stride = tgmax + ((9 - (tgmax & 7)) & 7);
        // Classify the tile's entries.  Culled: farther than the cutoff from
        // the pass's target box, every pair masked.  Unmasked: unit mass and
        // the farthest point of the box inside the cutoff, so every pair
        // passes the cutoff test and m_j = 1.  Masked: the rest.
        float4 v[NCH];
        uint64_t bk[NCH], ba[NCH];
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
          v[k] = pad;
          bk[k] = ba[k] = 0;
          if (k * 64 < nt) {
            const int q = k * 64 + l;
            bool keep = false, um = false;
            if (q < nt) {
              const PJ p = pj[jidx[t0 + q]];
              v[k] = make_float4(p.x, p.y, p.z, p.m);
              keep = !L.cull || box_d2(bmin, bmax, p.x, p.y, p.z) <= L.cull2;
              um = keep && L.unmasked && p.m == 1.f &&
                   box_far2(bmin, bmax, p.x, p.y, p.z) < L.in2;
            }
            bk[k] = __ballot(keep);
            ba[k] = __ballot(um);
          }
        }
        int nA = 0, nK = 0;
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
          nA += __popcll(ba[k]);
          nK += __popcll(bk[k]);
        }
        // Unmasked entries fill the first tA slots of every group (a multiple
        // of UNROLL, entry a -> group a % G, slot a / G); the unmasked ones
        // beyond G tA and the masked ones follow as masked entries m -> group
        // m % G, slot tA + m / G, padded to tM slots.
        const int tA = (nA >> lg) / UNROLL * UNROLL;
        const int splitA = tA << lg;
        const int nM = nK - splitA;
        const int tM = (((nM + G - 1) >> lg) + UNROLL - 1) / UNROLL * UNROLL;
        __syncthreads();
        const uint64_t lt = lanemask_lt();
        int pa = 0, pb = nA - splitA;
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
          const uint64_t bm = bk[k] & ~ba[k];
          if ((bk[k] >> l) & 1ull) {
            int slot, grp;
            if ((ba[k] >> l) & 1ull) {
              const int a = pa + __popcll(ba[k] & lt);
              const int m = a - splitA;
              grp = (a < splitA ? a : m) & (G - 1);
              slot = a < splitA ? (a >> lg) : tA + (m >> lg);
            } else {
              const int m = pb + __popcll(bm & lt);
              grp = m & (G - 1);
              slot = tA + (m >> lg);
            }
            tile[grp * stride + slot] = v[k];
          }
          pa += __popcll(ba[k]);
          pb += __popcll(bm);
        }
        for (int m = nM + l; m < (tM << lg); m += 64) tile[(m & (G - 1)) * stride + tA + (m >> lg)] = pad;
        ev += (int64_t)nK * ni;
        evu += (int64_t)splitA * ni;
        __syncthreads();
        if (active) {
          const float4 *tp = tile + g * stride;
          for (int k = 0; k < tA; k += UNROLL) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
              const float4 p = tp[k + u];
#pragma unroll
              for (int r = 0; r < R; ++r)
                pair<LAWK, ORDER, GUARD, false>(L, c0, rcs, p, xi[r], yi[r], zi[r], ax[r], ay[r], az[r]);
            }
          }
          for (int k = tA; k < tA + tM; k += UNROLL) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
              const float4 p = tp[k + u];
#pragma unroll
              for (int r = 0; r < R; ++r)
                pair<LAWK, ORDER, GUARD, true>(L, c0, rcs, p, xi[r], yi[r], zi[r], ax[r], ay[r], az[r]);
            }
          }
        }
      }
      // fixed-order reduction over the G groups
      __syncthreads();
      for (int q = l; q < 64 * R; q += 64) red[0][q] = red[1][q] = red[2][q] = 0.f;
      __syncthreads();
      for (int gg = 0; gg < G; ++gg) {
        if (active && g == gg) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int ii = tq + r * lpg;
            if (ii < ni) {
              red[0][ii] += ax[r];
              red[1][ii] += ay[r];
              red[2][ii] += az[r];
            }
          }
        }
        __syncthreads();
      }
      if (I.pbase < 0) {
        for (int ii = l; ii < ni; ii += 64) {
          const int32_t q = iperm[off + ib + ii];
          fx[q] = red[0][ii];
          fy[q] = red[1][ii];
          fz[q] = red[2][ii];
        }
      } else {
        float *pp = part + I.pbase;
        for (int ii = l; ii < ni; ii += 64) {
          pp[3 * (ib + ii) + 0] = red[0][ii];
          pp[3 * (ib + ii) + 1] = red[1][ii];
          pp[3 * (ib + ii) + 2] = red[2][ii];
        }
      }
    }
    if (l == 0 && ev > 0) atomicAdd(evals, (unsigned long long)ev);
    if (l == 0 && evu > 0) atomicAdd(evals + 1, (unsigned long long)evu);
  }
}

// Sum the chunk partials of long lists in chunk order.
__global__ void __launch_bounds__(64)
k_reduce_chunks(const hsr_node *cn, const int32_t *leaves,
                const int64_t *multi, const int64_t *nmulti_p,
                const int64_t *leaf_pbase, const int32_t *leaf_nchunk,
                const int32_t *iperm, const float *part, float *fx, float *fy, float *fz) {
  // a fixed grid over the device-resident count, so a captured force phase
  // serves every list build
  for (int64_t k = blockIdx.x; k < *nmulti_p; k += gridDim.x) {
    const int64_t lf = multi[k];
    const hsr_node &nd = cn[leaves[lf]];
    const int64_t cnt = nd.count, off = nd.offset;
    const int nc = leaf_nchunk[lf];
    const float *p0 = part + leaf_pbase[lf];
    for (int64_t i = threadIdx.x; i < cnt; i += 64) {
      float sx = 0.f, sy = 0.f, sz = 0.f;
      for (int c = 0; c < nc; ++c) {
        const float *p = p0 + (int64_t)c * 3 * cnt + 3 * i;
        sx += p[0];
        sy += p[1];
        sz += p[2];
      }
      const int32_t q = iperm[off + i];  // partials are kept in target order
      fx[q] = sx;
      fy[q] = sy;
      fz[q] = sz;
    }
  }
}

// Work items: chunk counts per leaf.
__global__ void k_item_count(int64_t nleaves, const int64_t *ent_ptr,
                             int64_t chunk, int32_t *nchunk, int64_t *nitem,
                             int64_t *multi_flag, const hsr_node *cn,
                             const int32_t *leaves, int64_t *psize) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  int64_t L = ent_ptr[k + 1] - ent_ptr[k];
  int64_t c = std::max<int64_t>(1, (L + chunk - 1) / chunk);
  nchunk[k] = (int32_t)c;
  nitem[k] = c;
  multi_flag[k] = c > 1 ? 1 : 0;
  psize[k] = c > 1 ? c * 3 * cn[leaves[k]].count : 0;
}

__global__ void k_item_fill(int64_t nleaves, const int64_t *ent_ptr,
                            int64_t chunk, const int32_t *nchunk,
                            const int64_t *item_base, const int64_t *pbase,
                            const hsr_node *cn, const int32_t *leaves,
                            Item *items, uint64_t *keys, int32_t *vals,
                            int64_t *multi, const int64_t *multi_pos,
                            const int64_t *multi_flag) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int64_t e0 = ent_ptr[k], e1 = ent_ptr[k + 1];
  const int nc = nchunk[k];
  const int64_t cnt = cn[leaves[k]].count;
  for (int c = 0; c < nc; ++c) {
    int64_t j = item_base[k] + c;
    Item I;
    I.leaf = (int32_t)k;
    I.chunk = c;
    I.e0 = e0 + (int64_t)c * chunk;
    I.e1 = std::min<int64_t>(e1, I.e0 + chunk);
    I.pbase = nc > 1 ? pbase[k] + (int64_t)c * 3 * cnt : -1;
    items[j] = I;
    // work ~ i-passes * entries
    uint64_t work = (uint64_t)((cnt + 64 * DEF_R - 1) / (64 * DEF_R)) * (uint64_t)(I.e1 - I.e0);
    keys[j] = work;
    vals[j] = (int32_t)j;
  }
  if (multi_flag[k]) multi[multi_pos[k]] = k;
}

__global__ void k_item_permute(int64_t n, const Item *in, const int32_t *order,
                               Item *out) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[order[k]];
}

__global__ void k_fill3(int64_t n, float *a, float *b, float *c) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = b[k] = c[k] = 0.f;
}

__global__ void k_zero_u64(unsigned long long *p) {
  if (threadIdx.x == 0) *p = 0ull;
}

__global__ void k_kick(int64_t nleaves, const hsr_node *cn,
                       const int32_t *leaves, const float *fx,
                       const float *fy, const float *fz, const float *ms,
                       float fcoeff, int mul_mass, float *vx, float *vy,
                       float *vz) {
  int64_t k = blockIdx.x;
  if (k >= nleaves) return;
  const hsr_node &nd = cn[leaves[k]];
  for (int64_t i = threadIdx.x; i < nd.count; i += blockDim.x) {
    int64_t p = nd.offset + i;
    float cfac = mul_mass ? fcoeff * ms[p] : fcoeff;
    vx[p] = vx[p] + fx[p] * cfac;
    vy[p] = vy[p] + fy[p] * cfac;
    vz[p] = vz[p] + fz[p] * cfac;
  }
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <class T>
hsr_status excl_scan(hsr_ctx *c, const T *in, T *out, int64_t n) {
  size_t tmp = 0;
  HSR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, c->stream));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp + 256));
  HSR_HIP(hipcub::DeviceScan::ExclusiveSum(c->s.cub.p, tmp, in, out, (int)n, c->stream));
  return HSR_OK;
}

// P(r2) = sum a_k r2^k re-expanded in s = r2 + eps2 (host, double).
LawArgs law_args(const ForceLawDev &D) {
  LawArgs A;
  double a[7] = {0, 0, 0, 0, 0, 0, 0}, b[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k <= 6; ++k) a[k] = D.c[k];
  const double e = D.eps2;
  // P(s - e) = sum_k a_k (s - e)^k
  for (int k = 0; k <= 6; ++k) {
    double binom = 1.0;
    for (int i = 0; i <= k; ++i) {
      // coefficient of s^i in (s - e)^k: C(k,i) (-e)^(k-i)
      double cki = binom;
      double term = cki * std::pow(-e, k - i);
      b[i] += a[k] * term;
      binom = binom * (double)(k - i) / (double)(i + 1);
    }
  }
  for (int k = 0; k <= 6; ++k) A.c[k] = (float)b[k];
  // r2 < rc2  <=>  s < rc2 + eps2 up to rounding at the boundary, where the
  // short-range force is ~1e-5 of its r = 1 value
  A.rcs = (float)((double)D.rc2 + e);
  A.eps2 = D.eps2;
  A.guard = D.guard;
  // a relative margin far above the few-ulp rounding of r2 keeps every pair
  // the kernel's own s < rcs test would accept
  A.cull = 1;
  A.cull2 = (float)((double)D.rc2 * (1.0 + 1e-4));
  A.pmax = 64 * DEF_R;
  // the same relative margin the other way: every pair of an entry whose
  // farthest box point is inside in2 passes the kernel's cutoff test
  A.in2 = (float)((double)D.rc2 * (1.0 - 1e-4));
  A.unmasked = D.guard ? 0 : 1;
  A.rc2 = D.rc2;
  // FGrid (ForceLaw.cxx:23-33); f0 = c + 2/3 b^3 as the reference rounds it
  A.b = 0.72f; A.cc = 0.01f; A.d = 0.27f; A.e = 0.0001f; A.ff = 360.0f; A.g = 100.0f;
  A.h = 0.67f; A.l = 17.0f; A.rmax = 3.116326355f;
  A.f0 = (float)(A.cc + 2.0 / 3.0 * A.b * A.b * A.b);
  A.bins = D.bins;
  A.t_r2min = D.t_r2min;
  A.t_r2max = D.t_r2max;
  A.t_oodr2 = D.t_oodr2;
  A.t_last = std::max(0, D.n_interp - 2);
  return A;
}

}  // namespace

// Work items of the current lists (rebuilt after every hsr_lists call).
// Target order inside each listed leaf: positions sorted along the leaf's
// longest box edge (ties x > y > z; equal coordinates by position), so that
// a pass's 128 targets span a slab of the leaf and the entries its box is
// beyond the cutoff of can be skipped.  iperm[off + k] = position of the
// k-th target; leaves above SORT_MAX keep their order.
constexpr int SORT_MAX = 1024;

__global__ void k_iota32(int64_t n, int32_t *a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = (int32_t)k;
}

__global__ void __launch_bounds__(256)
k_leaf_sort(int64_t nl, const int32_t *__restrict__ leaves, const hsr_node *__restrict__ cn,
            const PJ *__restrict__ pj, int32_t *__restrict__ iperm) {
  __shared__ float key[SORT_MAX];
  __shared__ int32_t idx[SORT_MAX];
  const int64_t k = blockIdx.x;
  if (k >= nl) return;
  const hsr_node &nd = cn[leaves[k]];
  const int64_t off = nd.offset, cnt = nd.count;
  if (cnt < 2 || cnt > SORT_MAX) return;
  const float e0 = nd.xmax[0] - nd.xmin[0], e1 = nd.xmax[1] - nd.xmin[1],
              e2 = nd.xmax[2] - nd.xmin[2];
  const int d = (e0 >= e1 && e0 >= e2) ? 0 : (e1 >= e2 ? 1 : 2);
  int np2 = 2;
  while (np2 < cnt) np2 <<= 1;
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    if (i < cnt) {
      const PJ p = pj[off + i];
      key[i] = d == 0 ? p.x : (d == 1 ? p.y : p.z);
    } else {
      key[i] = INFINITY;
    }
    idx[i] = i;
  }
  __syncthreads();
  // bitonic sort of (key, idx), ascending and lexicographic
  for (int size = 2; size <= np2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const float ki = key[i], kj = key[j];
          const int32_t ii = idx[i], ij = idx[j];
          const bool gt = ki > kj || (ki == kj && ii > ij);
          if (gt == up) {
            key[i] = kj;
            key[j] = ki;
            idx[i] = ij;
            idx[j] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) iperm[off + i] = (int32_t)(off + idx[i]);
}

// Shapes built for every polynomial law without the r2 > 0 guard (the
// production laws); the other laws run the default shape only.
struct Shape {
  int r, u, t;
};
constexpr Shape kShapes[] = {{2, 2, 256}, {2, 1, 256}, {2, 4, 256}, {1, 2, 256},
                             {1, 4, 256}, {4, 1, 256}, {2, 2, 128}, {2, 2, 512}};

void force_shape(const hsr_tuning &t, int *r, int *u, int *tile) {
  *r = t.i_per_lane ? t.i_per_lane : DEF_R;
  *u = t.unroll ? t.unroll : DEF_U;
  *tile = t.tile ? t.tile : DEF_T;
}

bool force_shape_built(int law_kind, bool guard, int r, int u, int tile) {
  if (r == DEF_R && u == DEF_U && tile == DEF_T) return true;
  if (law_kind != LAWK_POLY || guard) return false;
  for (const Shape &sh : kShapes)
    if (sh.r == r && sh.u == u && sh.t == tile) return true;
  return false;
}

std::string force_shapes_built() {
  std::string out;
  for (const Shape &sh : kShapes)
    out += (out.empty() ? "" : ", ") + std::string("(") + std::to_string(sh.r) + ", " +
           std::to_string(sh.u) + ", " + std::to_string(sh.t) + ")";
  return out;
}

hsr_status force_plan(hsr_ctx *c) {
  hipStream_t st = c->stream;
  ListState &LS = c->lists;
  const int64_t nl = LS.nleaves;
  const int64_t chunk = c->tuning.chunk > 0 ? c->tuning.chunk : 8192;
  LS.chunk = chunk;
  if (nl == 0) {
    LS.nitems = 0;
    return HSR_OK;
  }
  // scratch layout (int64 unless noted), arrays at a stride that only grows,
  // so their addresses (held by a captured force phase) survive list builds
  if (nl + 1 > LS.plan_stride) LS.plan_stride = std::max<int64_t>(nl + 1, LS.plan_stride + LS.plan_stride / 4);
  const int64_t ps = LS.plan_stride;
  HSR_HIP(dbuf_reserve(LS.plan, sizeof(int64_t) * 8 * ps + sizeof(int32_t) * ps + 256));
  int64_t *nitem = dptr<int64_t>(LS.plan);
  int64_t *item_base = nitem + ps;
  int64_t *mflag = item_base + ps;
  int64_t *mpos = mflag + ps;
  int64_t *psize = mpos + ps;
  int64_t *pbase = psize + ps;
  int64_t *multi = pbase + ps;
  int64_t *counts = multi + ps;  // [0] nitems, [1] nmulti, [2] partial floats
  int32_t *nchunk = reinterpret_cast<int32_t *>(counts + ps);
  const hsr_node *cn = dptr<hsr_node>(c->tree.canon_nodes);
  const int32_t *leaves = dptr<int32_t>(LS.leaves);
  const int64_t *ent_ptr = dptr<int64_t>(LS.ent_ptr);
  k_item_count<<<nblk(nl, 256), 256, 0, st>>>(nl, ent_ptr, chunk, nchunk, nitem, mflag, cn, leaves, psize);
  hsr_status s;
  if ((s = excl_scan(c, nitem, item_base, nl)) != HSR_OK) return s;
  if ((s = excl_scan(c, mflag, mpos, nl)) != HSR_OK) return s;
  if ((s = excl_scan(c, psize, pbase, nl)) != HSR_OK) return s;
  int64_t tail[6];
  HSR_HIP(hipMemcpyAsync(&tail[0], item_base + nl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipMemcpyAsync(&tail[1], nitem + nl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipMemcpyAsync(&tail[2], mpos + nl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipMemcpyAsync(&tail[3], mflag + nl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipMemcpyAsync(&tail[4], pbase + nl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipMemcpyAsync(&tail[5], psize + nl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(stream_wait(st));
  const int64_t nitems = tail[0] + tail[1];
  const int64_t nmulti = tail[2] + tail[3];
  const int64_t npart = tail[4] + tail[5];
  // the item sort and the item ids are 32-bit
  if (nitems >= (int64_t)INT32_MAX) {
    hsr_set_error("force plan: " + std::to_string(nitems) +
                  " work items (2^31 or more); raise tuning.chunk");
    return HSR_ERR_NOMEM;
  }
  const int64_t plan_bytes = (int64_t)sizeof(Item) * 2 * (nitems + 1) +
                             (int64_t)(sizeof(uint64_t) + sizeof(int32_t)) * 2 * (nitems + 1) +
                             (int64_t)sizeof(float) * npart;
  if (c->tuning.max_plan_bytes > 0 && plan_bytes > c->tuning.max_plan_bytes) {
    hsr_set_error("force plan: needs " + std::to_string(plan_bytes) + " bytes, above tuning.max_plan_bytes (" +
                  std::to_string(c->tuning.max_plan_bytes) + ")");
    return HSR_ERR_NOMEM;
  }
  LS.nitems = nitems;
  LS.nmulti = nmulti;
  HSR_HIP(dbuf_reserve(LS.items, sizeof(Item) * 2 * (nitems + 1)));
  HSR_HIP(dbuf_reserve(LS.item_keys, (sizeof(uint64_t) + sizeof(int32_t)) * 2 * (nitems + 1) + 256));
  HSR_HIP(dbuf_reserve(LS.part, sizeof(float) * std::max<int64_t>(npart, 1)));
  Item *items_tmp = dptr<Item>(LS.items) + (nitems + 1);
  Item *items = dptr<Item>(LS.items);
  uint64_t *keys = dptr<uint64_t>(LS.item_keys);
  uint64_t *keys2 = keys + (nitems + 1);
  int32_t *vals = reinterpret_cast<int32_t *>(keys2 + (nitems + 1));
  int32_t *vals2 = vals + (nitems + 1);
  k_item_fill<<<nblk(nl, 256), 256, 0, st>>>(nl, ent_ptr, chunk, nchunk, item_base, pbase, cn, leaves,
                                             items_tmp, keys, vals, multi, mpos, mflag);
  // heaviest items first
  size_t tmp = 0;
  HSR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, keys, keys2, vals, vals2, (int)nitems, 0, 64, st));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp + 256));
  HSR_HIP(hipcub::DeviceRadixSort::SortPairsDescending(c->s.cub.p, tmp, keys, keys2, vals, vals2, (int)nitems, 0, 64, st));
  k_item_permute<<<nblk(nitems, 256), 256, 0, st>>>(nitems, items_tmp, vals2, items);
  // device-resident counts for the persistent launch
  int64_t hc[2] = {nitems, nmulti};
  HSR_HIP(hipMemcpyAsync(counts, hc, sizeof(hc), hipMemcpyHostToDevice, st));
  HSR_HIP(stream_wait(st));
  HSR_HIP(dbuf_reserve(c->s.qctr, 64));
  LS.multi = multi;
  LS.counts = counts;
  LS.leaf_pbase = pbase;
  LS.leaf_nchunk = nchunk;
  // target order inside the listed leaves
  const int64_t n = c->tree.n;
  HSR_HIP(dbuf_reserve(LS.iperm, sizeof(int32_t) * std::max<int64_t>(n, 1)));
  int32_t *iperm = dptr<int32_t>(LS.iperm);
  if (n) k_iota32<<<nblk(n, 256), 256, 0, st>>>(n, iperm);
  if (!c->tuning.tree_order_targets)
    k_leaf_sort<<<(unsigned)nl, 256, 0, st>>>(nl, leaves, cn, dptr<PJ>(c->tree.pj), iperm);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

static hsr_status force_enqueue(hsr_ctx *c, float *fx, float *fy, float *fz,
                                bool timed);

static unsigned reduce_grid(const hsr_ctx *c) {
  unsigned g = 256;
  while ((int64_t)g < c->lists.nmulti) g <<= 1;
  return g;
}

// The force phase: zero F, reset the work queue, persistent leaf kernel,
// chunk reduction.  With tuning.use_graph the launches are captured once and
// replayed as one graph across list builds: their grids are fixed (the
// persistent kernel and the chunk reduction read their counts from device
// memory), so a new capture is needed only when an address they use, the
// particle count or the settings change.
hsr_status force_run(hsr_ctx *c, float *fx, float *fy, float *fz) {
  hipStream_t st = c->stream;
  if (c->law.kind == LAWK_INTERP && !c->law.bins) {
    hsr_set_error("HSR_LAW_SR_INTERP: no table (call hsr_set_law_table first)");
    return HSR_ERR_STATE;
  }
  if (!c->tuning.use_graph) return force_enqueue(c, fx, fy, fz, true);
  const ListState &LS = c->lists;
  const uintptr_t sig[20] = {(uintptr_t)fx, (uintptr_t)fy, (uintptr_t)fz, (uintptr_t)c->tree.n,
                             (uintptr_t)(LS.nleaves == 0 || LS.nitems == 0), (uintptr_t)c->tree.canon_nodes.p,
                             (uintptr_t)LS.leaves.p, (uintptr_t)LS.items.p, (uintptr_t)LS.counts,
                             (uintptr_t)LS.jidx.p, (uintptr_t)c->tree.pj.p, (uintptr_t)LS.iperm.p,
                             (uintptr_t)LS.part.p, (uintptr_t)LS.multi, (uintptr_t)LS.leaf_pbase,
                             (uintptr_t)LS.leaf_nchunk, (uintptr_t)c->s.qctr.p, (uintptr_t)reduce_grid(c), 0, 0};
  bool same = c->graph_exec != nullptr && c->graph_gen != -1;
  for (int k = 0; k < 20 && same; ++k) same = c->graph_sig[k] == sig[k];
  const bool stale = !same;
  if (stale) {
    if (c->graph_exec) {
      HSR_HIP(hipGraphExecDestroy(c->graph_exec));
      c->graph_exec = nullptr;
    }
    if (c->graph) {
      HSR_HIP(hipGraphDestroy(c->graph));
      c->graph = nullptr;
    }
    HSR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hsr_status s = force_enqueue(c, fx, fy, fz, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (s != HSR_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    HSR_HIP(e);
    c->graph = g;
    HSR_HIP(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
    c->graph_gen = 0;
    for (int k = 0; k < 20; ++k) c->graph_sig[k] = sig[k];
    c->graph_captures++;
  }
  HSR_HIP(hipEventRecord(c->ev[6], st));
  HSR_HIP(hipGraphLaunch(c->graph_exec, st));
  HSR_HIP(hipEventRecord(c->ev[7], st));
  c->kernel_timed = true;
  c->graph_launches++;
  return HSR_OK;
}

static hsr_status force_enqueue(hsr_ctx *c, float *fx, float *fy, float *fz,
                                bool timed) {
  hipStream_t st = c->stream;
  const int64_t n = c->tree.n;
  ListState &LS = c->lists;
  if (timed) c->kernel_timed = false;
  if (n > 0)
    k_fill3<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, fx, fy, fz);
  if (LS.nleaves == 0 || LS.nitems == 0) return HSR_OK;
  unsigned long long *qctr = dptr<unsigned long long>(c->s.qctr);
  k_zero_u64<<<1, 64, 0, st>>>(qctr);
  k_zero_u64<<<1, 64, 0, st>>>(qctr + 1);
  k_zero_u64<<<1, 64, 0, st>>>(qctr + 2);
  LawArgs A = law_args(c->law);
  A.cull = c->tuning.cull_off ? 0 : 1;
  int R, U, T;
  force_shape(c->tuning, &R, &U, &T);
  A.pmax = c->tuning.force_pass == 32 * R ? 32 * R : 64 * R;
  const int wpc = c->tuning.waves_per_cu > 0 ? c->tuning.waves_per_cu : 32;
  // the persistent grid does not depend on the item count (waves beyond the
  // items leave at their first pop), so a captured launch fits any list
  const int64_t grid = (int64_t)std::max(1, c->num_cus) * wpc;
  const hsr_node *cn = dptr<hsr_node>(c->tree.canon_nodes);
  const int32_t *leaves = dptr<int32_t>(LS.leaves);
  const Item *items = dptr<Item>(LS.items);
  float *part = dptr<float>(LS.part);
  // events bracket the leaf-kernel launch alone (the roofline's duration)
  if (timed) HSR_HIP(hipEventRecord(c->ev[6], st));
#define LAUNCH(LK, ORD, GRD, RR, UU, TT)                                              \
  k_force<LK, ORD, GRD, RR, UU, TT><<<(unsigned)grid, 64, 0, st>>>(cn, leaves, items, LS.counts, \
      dptr<int32_t>(LS.jidx), dptr<PJ>(c->tree.pj), dptr<int32_t>(LS.iperm), A, qctr, fx, fy, fz, \
      part, qctr + 1)
#define SHAPES(ORD)                                                                      \
  if (R == 2 && U == 1 && T == 256) LAUNCH(LAWK_POLY, ORD, false, 2, 1, 256);            \
  else if (R == 2 && U == 4 && T == 256) LAUNCH(LAWK_POLY, ORD, false, 2, 4, 256);       \
  else if (R == 1 && U == 2 && T == 256) LAUNCH(LAWK_POLY, ORD, false, 1, 2, 256);       \
  else if (R == 1 && U == 4 && T == 256) LAUNCH(LAWK_POLY, ORD, false, 1, 4, 256);       \
  else if (R == 4 && U == 1 && T == 256) LAUNCH(LAWK_POLY, ORD, false, 4, 1, 256);       \
  else if (R == 2 && U == 2 && T == 128) LAUNCH(LAWK_POLY, ORD, false, 2, 2, 128);       \
  else if (R == 2 && U == 2 && T == 512) LAUNCH(LAWK_POLY, ORD, false, 2, 2, 512);       \
  else LAUNCH(LAWK_POLY, ORD, false, DEF_R, DEF_U, DEF_T)
  if (c->law.kind == LAWK_FIT) {
    if (A.guard) LAUNCH(LAWK_FIT, 0, true, DEF_R, DEF_U, DEF_T);
    else LAUNCH(LAWK_FIT, 0, false, DEF_R, DEF_U, DEF_T);
  } else if (c->law.kind == LAWK_INTERP) {
    if (A.guard) LAUNCH(LAWK_INTERP, 0, true, DEF_R, DEF_U, DEF_T);
    else LAUNCH(LAWK_INTERP, 0, false, DEF_R, DEF_U, DEF_T);
  } else if (c->law.order == 6) {
    if (A.guard) LAUNCH(LAWK_POLY, 6, true, DEF_R, DEF_U, DEF_T);
    else { SHAPES(6); }
  } else {
    if (A.guard) LAUNCH(LAWK_POLY, 5, true, DEF_R, DEF_U, DEF_T);
    else { SHAPES(5); }
  }
#undef SHAPES
#undef LAUNCH
  HSR_HIP(hipGetLastError());
  if (timed) {
    HSR_HIP(hipEventRecord(c->ev[7], st));
    c->kernel_timed = true;
  }
  // a grid over the device-resident multi-chunk count (which may be 0): a
  // wave per multi-chunk leaf, the grid rounded up to a power of two so a
  // captured force phase serves later lists unless their count outgrows it
  k_reduce_chunks<<<reduce_grid(c), 64, 0, st>>>(
      cn, leaves, LS.multi, LS.counts + 1, LS.leaf_pbase, LS.leaf_nchunk, dptr<int32_t>(LS.iperm), part,
      fx, fy, fz);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

hsr_status kick_run(hsr_ctx *c, float fcoeff, float *vx, float *vy, float *vz,
                    const float *mass_sorted) {
  int64_t n = c->tree.n;
  HSR_HIP(dbuf_reserve(c->force, sizeof(float) * 3 * std::max<int64_t>(n, 1)));
  float *f = dptr<float>(c->force);
  hsr_status s = force_run(c, f, f + n, f + 2 * n);
  if (s != HSR_OK) return s;
  if (c->lists.nleaves == 0) return HSR_OK;
  // Step16 path: v += fcoeff * a (BGQStep16 caller, RCBForceTree.cxx:594-596);
  // ForceLaw path: nbody1 multiplies by m_i (RCBForceTree.cxx:610).
  int mul_mass = c->p.law == HSR_LAW_STEP16 ? 0 : 1;
  k_kick<<<(unsigned)c->lists.nleaves, 256, 0, c->stream>>>(
      c->lists.nleaves, dptr<hsr_node>(c->tree.canon_nodes),
      dptr<int32_t>(c->lists.leaves), f, f + n, f + 2 * n, mass_sorted, fcoeff,
      mul_mass, vx, vy, vz);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}
