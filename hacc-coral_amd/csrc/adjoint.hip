// adjoint.hip -- vector-Jacobian product of the short-range force with
// respect to positions (no counterpart in the reference; derived from the
// pair law of BGQStep16.c:167-187 / nbody1, RCBForceTree.cxx:599-619).
//
//   F_i = sum_j m_j g(s_ij) d_ij,  d_ij = x_j - x_i,  s = |d|^2 + rsm^2,
//   g(s) = [s^-3/2 - P(s)] [s < rc_s],  g'(s) = [-3/2 s^-5/2 - P'(s)] [s < rc_s]
//   J = sum_i lambda_i . F_i
//   dJ/dx_j +=  m_j [g lambda_i + 2 g' (d . lambda_i) d]     (j a source)
//   dJ/dx_i += -m_j [g lambda_i + 2 g' (d . lambda_i) d]     (i a target)
//
// A pseudo-particle sits at its node's centre of mass xc_n = sum m_k x_k / M_n
// (TDPTS = 1, RCBForceTree.cxx:525-569), so its source gradient G_n reaches
// each member k as (m_k / M_n) G_n.  The tree topology and lists are frozen
// (the cutoff and opening decisions are a measure-zero set).
//
// Two schedules, both deterministic (no float atomics):
//   fused (default)  one persistent kernel over the forward list items
//                 evaluates each pair once: the target side accumulates in
//                 registers, the source side W_e = sum_i V_ie of every list
//                 entry e is summed over the lanes holding i (DPP) and stored
//                 per entry; a second kernel sums W over the entries that
//                 reference each source particle in the order of a stable sort
//                 of the list segments by source, and pseudo-particle entries
//                 are summed per node and pushed down to the members;
//   two-pass      target pass as above without W; source pass per source leaf
//                 over the transposed lists (all target particles whose lists
//                 hold the leaf); pseudo pass per (target leaf, pseudo
//                 segment).  Used when W does not fit in device memory.
#include "hsr_internal.h"
#include "wave.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int R = 2;
constexpr int TILE = 256;
constexpr int GMAX = 16;
constexpr int LDS_ENT = 336;  // see force.hip: max G * stride for TILE 256
constexpr int UNROLL = 2;

struct AItem {
  int32_t leaf;     // node id of the accumulating leaf (canonical, global)
  int32_t pad;
  int64_t e0, e1;   // entry range in the entry array
  int64_t pbase;    // partial offset (-1: direct)
};

struct AdjLaw {
  float c[7];   // P(s) coefficients
  float dp[6];  // 2 * k * c_k: 2 P'(s) coefficients (dp[k-1] for s^(k-1))
  float rcs, eps2;
  int order;
  int cull;     // drop entries farther than the cutoff from a pass's target box
  float cull2;  // squared box distance beyond which every pair is masked
};

// Squared distance from p to the box [lo, hi] (0 inside).
__device__ __forceinline__ float box_d2(const float *lo, const float *hi, float x, float y,
                                        float z) {
  const float dx = fmaxf(0.f, fmaxf(lo[0] - x, x - hi[0]));
  const float dy = fmaxf(0.f, fmaxf(lo[1] - y, y - hi[1]));
  const float dz = fmaxf(0.f, fmaxf(lo[2] - z, z - hi[2]));
  return dx * dx + dy * dy + dz * dz;
}

AdjLaw adj_law(const ForceLawDev &D) {
  AdjLaw A;
  double a[7] = {0, 0, 0, 0, 0, 0, 0}, b[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k <= 6; ++k) a[k] = D.c[k];
  const double e = D.eps2;
  for (int k = 0; k <= 6; ++k) {
    double binom = 1.0;
    for (int i = 0; i <= k; ++i) {
      b[i] += a[k] * binom * std::pow(-e, k - i);
      binom = binom * (double)(k - i) / (double)(i + 1);
    }
  }
  for (int k = 0; k <= 6; ++k) A.c[k] = (float)b[k];
  for (int k = 1; k <= 6; ++k) A.dp[k - 1] = (float)(2.0 * k * b[k]);
  A.rcs = (float)((double)D.rc2 + e);
  A.eps2 = D.eps2;
  A.order = D.order;
  A.cull = 1;
  A.cull2 = (float)((double)D.rc2 * (1.0 + 1e-4));
  return A;
}

// g and 2 g' at s (masked by the cutoff), evaluated in s.
template <int ORDER>
__device__ __forceinline__ void pair_g(const AdjLaw &L, float s, float &g,
                                       float &g2p) {
  const float y = __builtin_amdgcn_rsqf(s);
  const float y2 = y * y;
  const float y3 = y2 * y;
  const float y5 = y3 * y2;
  float q, p;
  if (ORDER == 6) {
    q = fmaf(L.c[6], s, L.c[5]);
    q = fmaf(q, s, L.c[4]);
    p = fmaf(L.dp[5], s, L.dp[4]);
    p = fmaf(p, s, L.dp[3]);
  } else {
    q = fmaf(L.c[5], s, L.c[4]);
    p = fmaf(L.dp[4], s, L.dp[3]);
  }
  q = fmaf(q, s, L.c[3]);
  q = fmaf(q, s, L.c[2]);
  q = fmaf(q, s, L.c[1]);
  p = fmaf(p, s, L.dp[2]);
  p = fmaf(p, s, L.dp[1]);
  p = fmaf(p, s, L.dp[0]);
  g = fmaf(-s, q, y3 - L.c[0]);
  g2p = fmaf(-3.0f, y5, -p);
}

// MODE 0 (target pass): lanes hold targets i with lambda_i; staged entries
//   are sources (x_j, m_j); acc -= m_j [g lam_i + 2g'(d.lam_i) d], d = x_j - x_i.
// MODE 1 (source pass): lanes hold sources k; staged entries are targets
//   (x_i, lambda_i); acc += [g lam_i + 2g'(d.lam_i) d], d = x_k - x_i; the
//   result is scaled by m_k at the end.
template <int MODE, int ORDER>
__global__ void __launch_bounds__(64)
k_adj(const hsr_node *__restrict__ cn, const AItem *__restrict__ items,
      const int64_t *__restrict__ nitems_p, const int32_t *__restrict__ eidx,
      const PJ *__restrict__ pj, const float *__restrict__ lx,
      const float *__restrict__ ly, const float *__restrict__ lz, AdjLaw L,
      unsigned long long *__restrict__ qctr, float *__restrict__ gx,
      float *__restrict__ gy, float *__restrict__ gz,
      float *__restrict__ part) {
  __shared__ float4 tA[LDS_ENT];
  __shared__ float4 tB[LDS_ENT];
  float (*red)[64 * R] = reinterpret_cast<float (*)[64 * R]>(&tA[0]);
  const int l = lane_id();
  const int64_t nitems = *nitems_p;
  for (;;) {
    int64_t it = 0;
    if (l == 0) it = (int64_t)atomicAdd(qctr, 1ull);
    it = (int64_t)readlane_u64((uint64_t)it, 0);
    if (it >= nitems) break;
    const AItem I = items[it];
    const hsr_node &nd = cn[I.leaf];
    const int64_t off = nd.offset, cnt = nd.count;
    const float4 padA = make_float4(nd.xc[0] + 1000.f, nd.xc[1], nd.xc[2], 0.f);
    for (int64_t ib = 0; ib < cnt; ib += 64 * R) {
      const int ni = (int)min<int64_t>(64 * R, cnt - ib);
      const int lpg = max((ni + R - 1) / R, 64 / GMAX);
      const int G = max(1, 64 / lpg);
      const int g = l / lpg, tq = l - g * lpg;
      const bool active = g < G;
      float xi[R], yi[R], zi[R], li[R][3], ax[R], ay[R], az[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int ii = tq + r * lpg;
        const bool v = active && ii < ni;
        const int64_t p = off + ib + (v ? ii : 0);
        const PJ q = pj[p];
        xi[r] = v ? q.x : 1e30f;
        yi[r] = v ? q.y : 1e30f;
        zi[r] = v ? q.z : 1e30f;
        if (MODE == 0) {
          li[r][0] = v ? lx[p] : 0.f;
          li[r][1] = v ? ly[p] : 0.f;
          li[r][2] = v ? lz[p] : 0.f;
        }
        ax[r] = ay[r] = az[r] = 0.f;
      }
      for (int64_t t0 = I.e0; t0 < I.e1; t0 += TILE) {
        const int nt = (int)min<int64_t>(TILE, I.e1 - t0);
        int tg = (nt + G - 1) / G;
        tg = (tg + 3) / 4 * 4;
        const int stride = tg + ((9 - (tg & 7)) & 7);
        __syncthreads();
        for (int q = l; q < G * tg; q += 64) {
          const int gq = q / tg, kq = q - gq * tg;
          float4 a = padA, b = make_float4(0.f, 0.f, 0.f, 0.f);
          if (q < nt) {
            const int32_t j = eidx[t0 + q];
            const PJ p = pj[j];
            if (MODE == 0) {
              a = make_float4(p.x, p.y, p.z, p.m);
            } else {
              a = make_float4(p.x, p.y, p.z, lx[j]);
              b = make_float4(ly[j], lz[j], 0.f, 0.f);
            }
          }
          tA[gq * stride + kq] = a;
          if (MODE == 1) tB[gq * stride + kq] = b;
        }
        __syncthreads();
        if (active) {
          const float4 *pa = tA + g * stride;
          const float4 *pb = tB + g * stride;
          for (int k = 0; k < tg; k += UNROLL) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
              const float4 a = pa[k + u];
              float4 b;
              if (MODE == 1) b = pb[k + u];
#pragma unroll
              for (int r = 0; r < R; ++r) {
                float dx, dy, dz, lax, lay, laz;
                if (MODE == 0) {
                  dx = a.x - xi[r];
                  dy = a.y - yi[r];
                  dz = a.z - zi[r];
                  lax = li[r][0];
                  lay = li[r][1];
                  laz = li[r][2];
                } else {
                  dx = xi[r] - a.x;
                  dy = yi[r] - a.y;
                  dz = zi[r] - a.z;
                  lax = a.w;
                  lay = b.x;
                  laz = b.y;
                }
                float s = fmaf(dx, dx, L.eps2);
                s = fmaf(dy, dy, s);
                s = fmaf(dz, dz, s);
                float gg, g2p;
                pair_g<ORDER>(L, s, gg, g2p);
                const bool in = s < L.rcs;
                float dl = dx * lax;
                dl = fmaf(dy, lay, dl);
                dl = fmaf(dz, laz, dl);
                float c1, c2;
                if (MODE == 0) {
                  const float mm = in ? -a.w : 0.f;
                  c1 = mm * gg;
                  c2 = (mm * g2p) * dl;
                } else {
                  c1 = in ? gg : 0.f;
                  c2 = in ? g2p * dl : 0.f;
                }
                ax[r] = fmaf(c1, lax, ax[r]);
                ay[r] = fmaf(c1, lay, ay[r]);
                az[r] = fmaf(c1, laz, az[r]);
                ax[r] = fmaf(c2, dx, ax[r]);
                ay[r] = fmaf(c2, dy, ay[r]);
                az[r] = fmaf(c2, dz, az[r]);
              }
            }
          }
        }
      }
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
      for (int ii = l; ii < ni; ii += 64) {
        float vx = red[0][ii], vy = red[1][ii], vz = red[2][ii];
        const int64_t p = off + ib + ii;
        if (MODE == 1) {
          const float mk = pj[p].m;
          vx *= mk;
          vy *= mk;
          vz *= mk;
        }
        if (I.pbase < 0) {
          if (MODE == 0) {
            gx[p] = vx;
            gy[p] = vy;
            gz[p] = vz;
          } else {
            gx[p] += vx;
            gy[p] += vy;
            gz[p] += vz;
          }
        } else {
          float *pp = part + I.pbase + 3 * (ib + ii);
          pp[0] = vx;
          pp[1] = vy;
          pp[2] = vz;
        }
      }
    }
  }
}

// Chunked items: partials summed in chunk order.  MODE 0 writes, MODE 1 adds.
__global__ void __launch_bounds__(64)
k_adj_reduce(const hsr_node *cn, const AItem *items, const int64_t *first,
             int64_t nmulti, int add, const int32_t *iperm, const float *part, float *gx,
             float *gy, float *gz) {
  const int64_t k = blockIdx.x;
  if (k >= nmulti) return;
  const int64_t i0 = first[k], i1 = first[k + 1];
  if (i1 - i0 <= 1) return;  // single-chunk items wrote their result directly
  const hsr_node &nd = cn[items[i0].leaf];
  for (int64_t i = threadIdx.x; i < nd.count; i += 64) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int64_t c = i0; c < i1; ++c) {
      const float *p = part + items[c].pbase + 3 * i;
      sx += p[0];
      sy += p[1];
      sz += p[2];
    }
    // partials of the fused schedule are kept in target order
    const int64_t q = iperm ? (int64_t)iperm[nd.offset + i] : nd.offset + i;
    if (add) {
      gx[q] += sx;
      gy[q] += sy;
      gz[q] += sz;
    } else {
      gx[q] = sx;
      gy[q] = sy;
      gz[q] = sz;
    }
  }
}

// ------------------------------------------------------------- fused pass
constexpr int FR = 2;                       // targets per lane
#ifndef HSR_FUSED_TILE
#define HSR_FUSED_TILE 256
#endif
constexpr int FTILE = HSR_FUSED_TILE;       // entries per LDS tile
#ifndef HSR_FUSED_UNROLL
#define HSR_FUSED_UNROLL 1
#endif
constexpr int FUNROLL = HSR_FUSED_UNROLL;  // entries per step of a slice

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v) {
  const int t = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xf, true);
  return v + __builtin_bit_cast(float, t);
}

// Sum over aligned groups of LP lanes (a power of two, 4..64); the sum lands
// in the group's last lane.  Needs all 64 lanes active.
//   row_shr:1,2,4,8 sum within 16-lane rows; row_bcast:15 adds row 0 (2) into
//   row 1 (3); row_bcast:31 adds lane 31 into rows 2-3.
template <int LP>
__device__ __forceinline__ float group_sum_last(float v) {
  v = dpp_add<0x111, 0xf>(v);
  v = dpp_add<0x112, 0xf>(v);
  if (LP >= 8) v = dpp_add<0x114, 0xf>(v);
  if (LP >= 16) v = dpp_add<0x118, 0xf>(v);
  if (LP >= 32) v = dpp_add<0x142, 0xa>(v);
  if (LP >= 64) v = dpp_add<0x143, 0xc>(v);
  return v;
}

// Targets of the next pass over a leaf: 128 while that many remain, then the
// largest of 64, 32, 16, 8 that fits, then the rest (< 8).  With two targets
// per lane each chunk fills groups of a power-of-two width exactly.
__device__ __forceinline__ int64_t pass_size(int64_t rem) {
  if (rem >= 64 * FR) return 64 * FR;
  for (int64_t c = 32 * FR; c >= 4 * FR; c >>= 1)
    if (rem >= c) return c;
  return rem;
}

// A group's slice of the staged tile: every pair of its FR targets per lane
// with each entry, the target sums B in registers, and the entry's source sum
// W (over the group's lpg = LP lanes) written back into its tile slot.
template <int ORDER, int LP>
__device__ __forceinline__ void fused_slice(const AdjLaw &L, float4 *pa, int tg, bool last,
                                            const float (&xi)[FR], const float (&yi)[FR],
                                            const float (&zi)[FR], const float (&la)[FR],
                                            const float (&lb)[FR], const float (&lc)[FR],
                                            float (&am)[FR], float (&bx)[FR], float (&by)[FR],
                                            float (&bz)[FR]) {
  for (int k = 0; k < tg; k += FUNROLL) {
    float4 a[FUNROLL];
#pragma unroll
    for (int u = 0; u < FUNROLL; ++u) a[u] = pa[k + u];
    float vx[FUNROLL], vy[FUNROLL], vz[FUNROLL];
#pragma unroll
    for (int u = 0; u < FUNROLL; ++u) {
      vx[u] = vy[u] = vz[u] = 0.f;
#pragma unroll
      for (int r = 0; r < FR; ++r) {
        const float dx = a[u].x - xi[r];
        const float dy = a[u].y - yi[r];
        const float dz = a[u].z - zi[r];
        float s = fmaf(dx, dx, L.eps2);
        s = fmaf(dy, dy, s);
        s = fmaf(dz, dz, s);
        float gg, g2p;
        pair_g<ORDER>(L, s, gg, g2p);
        const bool in = s < L.rcs;
        float dl = dx * la[r];
        dl = fmaf(dy, lb[r], dl);
        dl = fmaf(dz, lc[r], dl);
        const float gm = in ? gg : 0.f;
        const float cc = in ? g2p * dl : 0.f;
        vx[u] = fmaf(gm, la[r], vx[u]);
        vy[u] = fmaf(gm, lb[r], vy[u]);
        vz[u] = fmaf(gm, lc[r], vz[u]);
        vx[u] = fmaf(cc, dx, vx[u]);
        vy[u] = fmaf(cc, dy, vy[u]);
        vz[u] = fmaf(cc, dz, vz[u]);
        // target: m_j V = (m_j g) lam_i + (m_j cc) d
        am[r] = fmaf(a[u].w, gm, am[r]);
        const float mc = a[u].w * cc;
        bx[r] = fmaf(mc, dx, bx[r]);
        by[r] = fmaf(mc, dy, by[r]);
        bz[r] = fmaf(mc, dz, bz[r]);
      }
    }
#pragma unroll
    for (int u = 0; u < FUNROLL; ++u) {
      vx[u] = group_sum_last<LP>(vx[u]);
      vy[u] = group_sum_last<LP>(vy[u]);
      vz[u] = group_sum_last<LP>(vz[u]);
    }
    // each slot was read above by every lane; its group's sum replaces it
    if (last) {
#pragma unroll
      for (int u = 0; u < FUNROLL; ++u) pa[k + u] = make_float4(vx[u], vy[u], vz[u], 0.f);
    }
  }
}

// One pair evaluation for both sides.  Lanes hold targets i (x_i, lambda_i)
// in groups of lpg lanes (FR per lane); each group takes a slice of the
// staged entries j (x_j, m_j).  Per pair, V = g lam_i + 2g'(d.lam_i) d:
//   target   A_i += m_j g, B_i += m_j 2g'(d.lam_i) d;  dJ/dx_i = -(A_i lam_i + B_i)
//   source   W_j += V, summed over the group's lanes (DPP) into the tile slot
//            the entry was staged in, then over the passes in W itself.
// Passes over a leaf's targets run one after the other (pass p > 0 adds to W;
// the barrier between passes orders those global writes within the group).
template <int ORDER>
__device__ __forceinline__ void
adj_fused_body(const hsr_node *__restrict__ cn, const AItem *__restrict__ items,
               const int64_t *__restrict__ nitems_p, const int32_t *__restrict__ jidx,
               const PJ *__restrict__ pj, const int32_t *__restrict__ iperm,
               const float *__restrict__ lx, const float *__restrict__ ly,
               const float *__restrict__ lz, const AdjLaw &L,
               unsigned long long *__restrict__ qctr, float *__restrict__ gx,
               float *__restrict__ gy, float *__restrict__ gz, float *__restrict__ part,
               float *__restrict__ wx, float *__restrict__ wy, float *__restrict__ wz) {
  __shared__ float4 tile[LDS_ENT];
  __shared__ int16_t qmap[FTILE];
  float (*red)[64 * FR] = reinterpret_cast<float (*)[64 * FR]>(&tile[0]);
  const int l = lane_id();
  const int64_t nitems = *nitems_p;
  for (;;) {
    int64_t it = 0;
    if (l == 0) it = (int64_t)atomicAdd(qctr, 1ull);
    it = (int64_t)readlane_u64((uint64_t)it, 0);
    if (it >= nitems) break;
    const AItem I = items[it];
    const hsr_node &nd = cn[I.leaf];
    const int64_t off = nd.offset, cnt = nd.count;
    if (cnt == 0) {  // an empty leaf contributes nothing to its entries
      for (int64_t e = I.e0 + l; e < I.e1; e += 64) wx[e] = wy[e] = wz[e] = 0.f;
      continue;
    }
    const float4 padJ = make_float4(nd.xc[0] + 1000.f, nd.xc[1], nd.xc[2], 0.f);
    const float padI = nd.xc[0] + 3000.f;
    // passes over the targets in power-of-two chunks (128, 64, ..., 8, rest)
    // so that every pass fills its groups of lp = chunk / FR lanes
    for (int64_t ib = 0, ni = 0; ib < cnt; ib += ni) {
      const bool first = ib == 0;
      ni = pass_size(cnt - ib);
      const int lp = max((int)(ni + FR - 1) / FR, 4);
      const int G = 64 / lp;
      const int g = l / lp, tq = l & (lp - 1);
      float xi[FR], yi[FR], zi[FR], la[FR], lb[FR], lc[FR], am[FR], bx[FR], by[FR], bz[FR];
      // the pass's targets (in the leaf's target order) and their box
      float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
      for (int r = 0; r < FR; ++r) {
        const int ii = tq + r * lp;
        const bool v = ii < ni;
        const int64_t q = iperm[off + ib + (v ? ii : 0)];
        const PJ P = pj[q];
        xi[r] = v ? P.x : padI;
        yi[r] = v ? P.y : nd.xc[1];
        zi[r] = v ? P.z : nd.xc[2];
        la[r] = v ? lx[q] : 0.f;
        lb[r] = v ? ly[q] : 0.f;
        lc[r] = v ? lz[q] : 0.f;
        am[r] = bx[r] = by[r] = bz[r] = 0.f;
        if (v) {
          bmin[0] = fminf(bmin[0], P.x); bmax[0] = fmaxf(bmax[0], P.x);
          bmin[1] = fminf(bmin[1], P.y); bmax[1] = fmaxf(bmax[1], P.y);
          bmin[2] = fminf(bmin[2], P.z); bmax[2] = fmaxf(bmax[2], P.z);
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
      for (int64_t t0 = I.e0; t0 < I.e1; t0 += FTILE) {
        const int nt = (int)min<int64_t>(FTILE, I.e1 - t0);
        int tgmax = (nt + G - 1) / G;
        tgmax = (tgmax + FUNROLL - 1) / FUNROLL * FUNROLL;
        const int stride = tgmax + ((9 - (tgmax & 7)) & 7);
        __syncthreads();
        // entries beyond the cutoff of the pass's targets add nothing to W in
        // this pass and are dropped (the first pass zeroes their W); the
        // kept ones go round-robin to the groups (c -> group c % G, slot
        // c / G), qmap remembering their place in the tile
        int nk = 0;
        for (int q0 = 0; q0 < nt; q0 += 64) {
          const int q = q0 + l;
          float4 a = padJ;
          bool keep = false;
          if (q < nt) {
            const PJ P = pj[jidx[t0 + q]];
            a = make_float4(P.x, P.y, P.z, P.m);
            keep = !L.cull || box_d2(bmin, bmax, P.x, P.y, P.z) <= L.cull2;
            if (!keep && first) wx[t0 + q] = wy[t0 + q] = wz[t0 + q] = 0.f;
          }
          const uint64_t bal = __ballot(keep);
          if (keep) {
            const int c = nk + __popcll(bal & lanemask_lt());
            tile[(c % G) * stride + c / G] = a;
            qmap[c] = (int16_t)q;
          }
          nk += __popcll(bal);
        }
        int tg = (nk + G - 1) / G;
        tg = (tg + FUNROLL - 1) / FUNROLL * FUNROLL;
        for (int c = nk + l; c < G * tg; c += 64) tile[(c % G) * stride + c / G] = padJ;
        __syncthreads();
        float4 *pa = tile + g * stride;
        const bool last = tq == lp - 1;
        switch (lp) {  // the group width fixes the DPP sequence
          case 4: fused_slice<ORDER, 4>(L, pa, tg, last, xi, yi, zi, la, lb, lc, am, bx, by, bz); break;
          case 8: fused_slice<ORDER, 8>(L, pa, tg, last, xi, yi, zi, la, lb, lc, am, bx, by, bz); break;
          case 16: fused_slice<ORDER, 16>(L, pa, tg, last, xi, yi, zi, la, lb, lc, am, bx, by, bz); break;
          case 32: fused_slice<ORDER, 32>(L, pa, tg, last, xi, yi, zi, la, lb, lc, am, bx, by, bz); break;
          default: fused_slice<ORDER, 64>(L, pa, tg, last, xi, yi, zi, la, lb, lc, am, bx, by, bz); break;
        }
        __syncthreads();
        for (int c = l; c < nk; c += 64) {
          const float4 w = tile[(c % G) * stride + c / G];
          const int64_t e = t0 + qmap[c];
          if (first) {
            wx[e] = w.x;
            wy[e] = w.y;
            wz[e] = w.z;
          } else {
            wx[e] += w.x;
            wy[e] += w.y;
            wz[e] += w.z;
          }
        }
      }
      // target side: sum the groups in order
      __syncthreads();
      for (int q = l; q < 64 * FR; q += 64) red[0][q] = red[1][q] = red[2][q] = red[3][q] = 0.f;
      __syncthreads();
      for (int gg = 0; gg < G; ++gg) {
        if (g == gg) {
#pragma unroll
          for (int r = 0; r < FR; ++r) {
            const int ii = tq + r * lp;
            if (ii < ni) {
              red[0][ii] += bx[r];
              red[1][ii] += by[r];
              red[2][ii] += bz[r];
              red[3][ii] += am[r];
            }
          }
        }
        __syncthreads();
      }
      for (int ii = l; ii < ni; ii += 64) {
        const int64_t q = iperm[off + ib + ii];
        const float a = red[3][ii];
        const float vx = -fmaf(a, lx[q], red[0][ii]);
        const float vy = -fmaf(a, ly[q], red[1][ii]);
        const float vz = -fmaf(a, lz[q], red[2][ii]);
        if (I.pbase < 0) {
          gx[q] = vx;
          gy[q] = vy;
          gz[q] = vz;
        } else {
          float *pp = part + I.pbase + 3 * (ib + ii);
          pp[0] = vx;
          pp[1] = vy;
          pp[2] = vz;
        }
      }
    }
  }
}

// Waves per SIMD the register allocation of the 5th-order body targets (0:
// the compiler's choice).
#ifndef HSR_ADJ_WAVES
#define HSR_ADJ_WAVES 5
#endif
#if HSR_ADJ_WAVES > 0
#define HSR_ADJ_ATTR __attribute__((amdgpu_waves_per_eu(HSR_ADJ_WAVES, HSR_ADJ_WAVES)))
#else
#define HSR_ADJ_ATTR
#endif
#define HSR_ADJ_FUSED_ARGS                                                             \
  const hsr_node *__restrict__ cn, const AItem *__restrict__ items,                    \
      const int64_t *__restrict__ nitems_p, const int32_t *__restrict__ jidx,          \
      const PJ *__restrict__ pj, const int32_t *__restrict__ iperm,                    \
      const float *__restrict__ lx,                                                    \
      const float *__restrict__ ly, const float *__restrict__ lz, AdjLaw L,            \
      unsigned long long *__restrict__ qctr, float *__restrict__ gx,                   \
      float *__restrict__ gy, float *__restrict__ gz, float *__restrict__ part,        \
      float *__restrict__ wx, float *__restrict__ wy, float *__restrict__ wz
__global__ void __launch_bounds__(64) HSR_ADJ_ATTR
k_adj_fused5(HSR_ADJ_FUSED_ARGS) {
  adj_fused_body<5>(cn, items, nitems_p, jidx, pj, iperm, lx, ly, lz, L, qctr, gx, gy, gz, part, wx, wy,
                    wz);
}
__global__ void __launch_bounds__(64)
k_adj_fused6(HSR_ADJ_FUSED_ARGS) {
  adj_fused_body<6>(cn, items, nitems_p, jidx, pj, iperm, lx, ly, lz, L, qctr, gx, gy, gz, part, wx, wy,
                    wz);
}

// Fused schedule, source side: one wave per source leaf (a run of sorted
// particle segments); particle k of the leaf receives m_k times the sum, over
// the run's segments in sorted order, of W at the segment's entry for k.
__global__ void __launch_bounds__(64)
k_src_reduce(int64_t nruns, int64_t nsp, const int64_t *run_first, const int32_t *key,
             const int32_t *val, const int32_t *seg_t, const int64_t *ent_ptr,
             const int64_t *seg_e, const hsr_node *cn, const PJ *pj, const float *wx,
             const float *wy, const float *wz, float *gx, float *gy, float *gz) {
  const int64_t r = blockIdx.x;
  if (r >= nruns) return;
  const int64_t p0 = run_first[r];
  const int64_t p1 = (r + 1 < nruns) ? run_first[r + 1] : nsp;
  const hsr_node &S = cn[key[p0]];
  for (int64_t k = threadIdx.x; k < S.count; k += 64) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int64_t p = p0; p < p1; ++p) {
      const int32_t sg = val[p];
      const int64_t e = ent_ptr[seg_t[sg]] + seg_e[sg] + k;
      sx += wx[e];
      sy += wy[e];
      sz += wz[e];
    }
    const int64_t q = S.offset + k;
    const float m = pj[q].m;
    gx[q] += m * sx;
    gy[q] += m * sy;
    gz[q] += m * sz;
  }
}

// Fused schedule, pseudo side: W of every sorted pseudo segment's entry.
__global__ void k_pseudo_from_w(int64_t np, int64_t p0, const int32_t *val,
                                const int32_t *seg_t, const int64_t *ent_ptr,
                                const int64_t *seg_e, const float *wx,
                                const float *wy, const float *wz, float *gseg) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int32_t sg = val[p0 + q];
  const int64_t e = ent_ptr[seg_t[sg]] + seg_e[sg];
  gseg[3 * q + 0] = wx[e];
  gseg[3 * q + 1] = wy[e];
  gseg[3 * q + 2] = wz[e];
}


// ---------------------------------------------------------- transposition
__global__ void k_seg_target(int64_t nl, const int64_t *seg_ptr,
                             int32_t *seg_t) {
  int64_t k = blockIdx.x;
  if (k >= nl) return;
  for (int64_t s = seg_ptr[k] + threadIdx.x; s < seg_ptr[k + 1]; s += blockDim.x)
    seg_t[s] = (int32_t)k;
}

__global__ void k_leaf_at(int64_t nn, const hsr_node *cn, int32_t *leaf_at) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nn) return;
  const hsr_node &o = cn[c];
  if (o.cl == 0 && o.cr == 0 && o.count > 0) leaf_at[o.offset] = (int32_t)c;
}

// Sort keys: particle segment -> source leaf id; pseudo segment -> nn + node;
// an empty segment (the self segment of an empty listed leaf) -> 2 nn + 1,
// after everything else (it has no source particle).
__global__ void k_seg_keys(int64_t ns, int64_t nn, const int64_t *sa,
                           const int64_t *sb, const int32_t *leaf_at,
                           int32_t *key, int32_t *val) {
  int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const int64_t b = sb[s];
  key[s] = b > 0 ? leaf_at[sa[s]] : (b < 0 ? (int32_t)(nn + sa[s]) : (int32_t)(2 * nn + 1));
  val[s] = (int32_t)s;
}

// Entry counts of the sorted particle segments (target leaf sizes), and run
// starts (first position of each key).
__global__ void k_tr_counts(int64_t ns, int64_t nn, const int32_t *key,
                            const int32_t *val, const int32_t *seg_t,
                            const hsr_node *cn, const int32_t *leaves,
                            int64_t *cnt, int32_t *start) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= ns) return;
  const int32_t k = key[p];
  cnt[p] = k < nn ? cn[leaves[seg_t[val[p]]]].count : 0;
  start[p] = (p == 0 || key[p - 1] != k) ? 1 : 0;
}

// One wave per sorted particle segment: entries = target leaf particles.
__global__ void __launch_bounds__(64)
k_tr_expand(int64_t nsp, const int32_t *val, const int32_t *seg_t,
            const hsr_node *cn, const int32_t *leaves, const int64_t *eoff,
            int32_t *tj) {
  int64_t p = blockIdx.x;
  if (p >= nsp) return;
  const hsr_node &t = cn[leaves[seg_t[val[p]]]];
  const int64_t e = eoff[p];
  for (int64_t i = threadIdx.x; i < t.count; i += 64) tj[e + i] = (int32_t)(t.offset + i);
}

// Source-pass items: one per run of particle segments (a source leaf),
// chunked; chunk partials laid out per item.
__global__ void k_tr_items(int64_t nruns, int64_t nsp, const int64_t *run_first,
                           const int32_t *key, const int64_t *eoff,
                           const int64_t *cnt, const hsr_node *cn,
                           int64_t chunk, int64_t *nchunks, int64_t *psize) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  const int64_t p0 = run_first[r];
  const int64_t p1 = (r + 1 < nruns) ? run_first[r + 1] : nsp;
  const int64_t e0 = eoff[p0], e1 = eoff[p1 - 1] + cnt[p1 - 1];
  const int64_t nc = std::max<int64_t>(1, (e1 - e0 + chunk - 1) / chunk);
  nchunks[r] = nc;
  // chunk partials of a chunked run: 3 floats per source-leaf particle per chunk
  psize[r] = nc > 1 ? nc * 3 * cn[key[p0]].count : 0;
}

__global__ void k_tr_items_fill(int64_t nruns, int64_t nsp,
                                const int64_t *run_first, const int32_t *key,
                                const int64_t *eoff, const int64_t *cnt,
                                const hsr_node *cn, int64_t chunk,
                                const int64_t *nchunks, const int64_t *ibase,
                                const int64_t *pbase, AItem *items,
                                int64_t *first) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  const int64_t p0 = run_first[r];
  const int64_t p1 = (r + 1 < nruns) ? run_first[r + 1] : nsp;
  const int64_t e0 = eoff[p0], e1 = eoff[p1 - 1] + cnt[p1 - 1];
  const int32_t leaf = key[p0];
  const int64_t nc = nchunks[r], lc = cn[leaf].count;
  for (int64_t c = 0; c < nc; ++c) {
    AItem I;
    I.leaf = leaf;
    I.pad = 0;
    I.e0 = e0 + c * chunk;
    I.e1 = std::min<int64_t>(e1, I.e0 + chunk);
    I.pbase = nc > 1 ? pbase[r] + c * 3 * lc : -1;
    items[ibase[r] + c] = I;
  }
  first[r] = ibase[r];
}

// Target-pass items from the forward work items (leaf index -> node id).
__global__ void k_fw_items(int64_t nl, const int32_t *leaves,
                           const int64_t *ent_ptr, const hsr_node *cn,
                           int64_t chunk, const int64_t *ibase,
                           const int64_t *pbase, AItem *items,
                           int64_t *first) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nl) return;
  const int64_t e0 = ent_ptr[k], e1 = ent_ptr[k + 1];
  const int64_t nc = std::max<int64_t>(1, (e1 - e0 + chunk - 1) / chunk);
  const int64_t lc = cn[leaves[k]].count;
  for (int64_t c = 0; c < nc; ++c) {
    AItem I;
    I.leaf = leaves[k];
    I.pad = 0;
    I.e0 = e0 + c * chunk;
    I.e1 = std::min<int64_t>(e1, I.e0 + chunk);
    I.pbase = nc > 1 ? pbase[k] + c * 3 * lc : -1;
    items[ibase[k] + c] = I;
  }
  first[k] = ibase[k];
}

__global__ void k_fw_counts(int64_t nl, const int64_t *ent_ptr,
                            const int32_t *leaves, const hsr_node *cn,
                            int64_t chunk, int64_t *nchunks, int64_t *psize) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nl) return;
  const int64_t L = ent_ptr[k + 1] - ent_ptr[k];
  const int64_t nc = std::max<int64_t>(1, (L + chunk - 1) / chunk);
  nchunks[k] = nc;
  psize[k] = nc > 1 ? nc * 3 * cn[leaves[k]].count : 0;
}

// Pseudo pass: one wave per sorted pseudo segment (node n listed by target
// leaf t): G = sum_{i in t} m_n [g lam_i + 2 g' (d.lam_i) d], d = p_n - x_i.
template <int ORDER>
__global__ void __launch_bounds__(64)
k_pseudo_grad(int64_t np, int64_t p0, int64_t nn, int64_t n,
              const int32_t *key, const int32_t *val, const int32_t *seg_t,
              const int32_t *leaves, const hsr_node *cn, const PJ *pj,
              const float *lx, const float *ly, const float *lz, AdjLaw L,
              float *gseg) {
  const int64_t q = blockIdx.x;
  if (q >= np) return;
  const int64_t p = p0 + q;
  const int64_t node = key[p] - nn;
  const PJ P = pj[n + node];
  const hsr_node &t = cn[leaves[seg_t[val[p]]]];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int64_t i = threadIdx.x; i < t.count; i += 64) {
    const int64_t a = t.offset + i;
    const PJ X = pj[a];
    const float dx = P.x - X.x, dy = P.y - X.y, dz = P.z - X.z;
    float s = fmaf(dx, dx, L.eps2);
    s = fmaf(dy, dy, s);
    s = fmaf(dz, dz, s);
    float gg, g2p;
    pair_g<ORDER>(L, s, gg, g2p);
    if (s < L.rcs) {
      const float la = lx[a], lb = ly[a], lc = lz[a];
      float dl = dx * la;
      dl = fmaf(dy, lb, dl);
      dl = fmaf(dz, lc, dl);
      const float c1 = P.m * gg, c2 = P.m * g2p * dl;
      sx = fmaf(c1, la, sx);
      sy = fmaf(c1, lb, sy);
      sz = fmaf(c1, lc, sz);
      sx = fmaf(c2, dx, sx);
      sy = fmaf(c2, dy, sy);
      sz = fmaf(c2, dz, sz);
    }
  }
  // fixed-order wave reduction (lane order)
  for (int o = 1; o < 64; o <<= 1) {
    float ox = __shfl_down(sx, o), oy = __shfl_down(sy, o), oz = __shfl_down(sz, o);
    if ((lane_id() & (2 * o - 1)) == 0) {
      sx += ox;
      sy += oy;
      sz += oz;
    }
  }
  if (lane_id() == 0) {
    gseg[3 * q + 0] = sx;
    gseg[3 * q + 1] = sy;
    gseg[3 * q + 2] = sz;
  }
}

// Per node: H_n = (sum of its pseudo segment terms, in order) / S_n, the push
// factor of its members.  S_n, the member mass sum over [offset, offset +
// count), is the denominator of the node's centre of mass xc (cm,
// BGQCM.c:181-212), so a member k moves the pseudo-particle by m_k / S_n.
// S_n differs from the pseudo mass ppm when the node holds an unsplittable
// leaf, whose ppm the reference builds from two empty child slots
// (RCBForceTree.cxx:727-729, 856-889) and leaves 0.  One wave per run of a
// node's sorted pseudo segments sums S_n in float, lanes strided over the
// members and then a fixed xor tree (deterministic; a float sum suffices for
// a derivative factor).  The two-pass schedule's terms already carry the
// pseudo mass ppm; the fused schedule's W sums do not, so they are scaled by
// ppm here.
__global__ void __launch_bounds__(64)
k_pseudo_node(int64_t np, int64_t p0, int64_t nn, const int32_t *key, const float *gseg,
              const hsr_node *cn, const PJ *pj, int with_ppm, float *H) {
  const int64_t q = blockIdx.x;
  if (q >= np) return;
  const int64_t p = p0 + q;
  if (q > 0 && key[p - 1] == key[p]) return;  // not a run start
  const int64_t node = key[p] - nn;
  const hsr_node &nd = cn[node];
  const int l = threadIdx.x;
  float m = 0.f;
  for (int64_t i = nd.offset + l; i < nd.offset + nd.count; i += 64) m += pj[i].m;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o);
  if (l != 0) return;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int64_t r = q; r < np && key[p0 + r] == key[p]; ++r) {
    sx += gseg[3 * r + 0];
    sy += gseg[3 * r + 1];
    sz += gseg[3 * r + 2];
  }
  const float f = m > 0.f ? (with_ppm ? nd.ppm[0] : 1.f) / m : 0.f;
  H[3 * node + 0] = sx * f;
  H[3 * node + 1] = sy * f;
  H[3 * node + 2] = sz * f;
}

__global__ void k_fill(int64_t n, float *a) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0.f;
}

__global__ void k_zero_u64a(unsigned long long *p) {
  if (threadIdx.x == 0) *p = 0ull;
}

// Canonical parent of every node.
__global__ void k_cparent(int64_t nn, const hsr_node *cn, int32_t *par) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nn) return;
  const hsr_node &o = cn[c];
  if (o.cl > 0) par[o.cl] = (int32_t)c;
  if (o.cr > 0) par[o.cr] = (int32_t)c;
}

__global__ void k_roots(int64_t nb, const int64_t *cbase, int32_t *par) {
  int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b < nb) par[cbase[b]] = -1;
}

// Members of every leaf receive m_k * sum over the leaf and its ancestors of
// G_n / M_n.
__global__ void __launch_bounds__(64)
k_pseudo_push(int64_t nn, const hsr_node *cn, const int32_t *par,
              const float *H, const PJ *pj, float *gx, float *gy, float *gz) {
  const int64_t c = blockIdx.x;
  if (c >= nn) return;
  const hsr_node &o = cn[c];
  if (o.cl != 0 || o.cr != 0 || o.count == 0) return;
  float hx = 0.f, hy = 0.f, hz = 0.f;
  for (int64_t a = c; a >= 0; a = par[a]) {
    hx += H[3 * a + 0];
    hy += H[3 * a + 1];
    hz += H[3 * a + 2];
  }
  if (hx == 0.f && hy == 0.f && hz == 0.f) return;
  for (int64_t i = threadIdx.x; i < o.count; i += 64) {
    const int64_t k = o.offset + i;
    const float m = pj[k].m;
    gx[k] += m * hx;
    gy[k] += m * hy;
    gz[k] += m * hz;
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

template <class T>
T tail_sum(hsr_ctx *c, const T *excl, const T *cnt, int64_t n) {
  T a = 0, b = 0;
  if (n == 0) return 0;
  (void)hipMemcpyAsync(&a, excl + n - 1, sizeof(T), hipMemcpyDeviceToHost, c->stream);
  (void)hipMemcpyAsync(&b, cnt + n - 1, sizeof(T), hipMemcpyDeviceToHost, c->stream);
  (void)hipStreamSynchronize(c->stream);
  return a + b;
}

}  // namespace

// Transposed lists and item queues (once per list build).
hsr_status adjoint_plan(hsr_ctx *c) {
  hipStream_t st = c->stream;
  AdjState &A = c->adj;
  ListState &LS = c->lists;
  TreeState &T = c->tree;
  const int64_t nl = LS.nleaves, ns = LS.nsegs, nn = T.nnodes, n = T.n;
  const int64_t chunk = LS.chunk > 0 ? LS.chunk : 16384;
  // host sources of asynchronous copies stay alive until the final sync
  int64_t tail_items[4] = {0, 0, 0, 0};
  const hsr_node *cn = dptr<hsr_node>(T.canon_nodes);
  const int32_t *leaves = dptr<int32_t>(LS.leaves);
  const int64_t *seg_ptr = dptr<int64_t>(LS.seg_ptr);
  const int64_t *sa = dptr<int64_t>(LS.seg);
  const int64_t *sb = sa + ns + 1;
  hsr_status s;
  // ---- target-pass items (forward entries)
  HSR_HIP(dbuf_reserve(A.fw, sizeof(int64_t) * 5 * (nl + 2)));
  int64_t *fnc = dptr<int64_t>(A.fw), *fib = fnc + (nl + 2), *fps = fib + (nl + 2);
  int64_t *fpb = fps + (nl + 2), *ffirst = fpb + (nl + 2);
  k_fw_counts<<<nblk(nl, 256), 256, 0, st>>>(nl, dptr<int64_t>(LS.ent_ptr), leaves, cn, chunk, fnc, fps);
  if ((s = excl_scan(c, fnc, fib, nl)) != HSR_OK) return s;
  if ((s = excl_scan(c, fps, fpb, nl)) != HSR_OK) return s;
  const int64_t nfi = tail_sum(c, fib, fnc, nl);
  const int64_t nfp = tail_sum(c, fpb, fps, nl);
  A.n_fw_items = nfi;
  A.n_fw_leaves = nl;
  HSR_HIP(dbuf_reserve(A.fw_items, sizeof(AItem) * (nfi + 1)));
  k_fw_items<<<nblk(nl, 256), 256, 0, st>>>(nl, leaves, dptr<int64_t>(LS.ent_ptr), cn, chunk, fib, fpb,
                                            dptr<AItem>(A.fw_items), ffirst);
  A.fw_first = ffirst;
  A.fw_part = nfp;
  // ---- transposition of the segments
  HSR_HIP(dbuf_reserve(A.seg, sizeof(int32_t) * 6 * (ns + 2) + sizeof(int64_t) * 3 * (ns + 2) +
                                  sizeof(int32_t) * (n + 1)));
  int32_t *seg_t = dptr<int32_t>(A.seg);
  int32_t *key = seg_t + (ns + 2), *val = key + (ns + 2), *key2 = val + (ns + 2), *val2 = key2 + (ns + 2);
  int32_t *start = val2 + (ns + 2);
  int64_t *tcnt = reinterpret_cast<int64_t *>(start + (ns + 2));
  int64_t *eoff = tcnt + (ns + 2);
  int64_t *run_first = eoff + (ns + 2);
  int32_t *leaf_at = reinterpret_cast<int32_t *>(run_first + (ns + 2));
  k_seg_target<<<(unsigned)std::max<int64_t>(nl, 1), 64, 0, st>>>(nl, seg_ptr, seg_t);
  k_leaf_at<<<nblk(nn, 256), 256, 0, st>>>(nn, cn, leaf_at);
  k_seg_keys<<<nblk(ns, 256), 256, 0, st>>>(ns, nn, sa, sb, leaf_at, key, val);
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) <= 2 * nn + 1) ++bits;
  size_t tmp = 0;
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key2, val, val2, (int)ns, 0, bits, st));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp + 256));
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(c->s.cub.p, tmp, key, key2, val, val2, (int)ns, 0, bits, st));
  A.key = key2;
  A.val = val2;
  A.seg_t = seg_t;
  // ---- schedule: fused unless W (12 B per list entry) does not fit
  {
    const size_t wbytes = sizeof(float) * 3 * (size_t)(LS.nentries + 1);
    size_t free_b = 0, total_b = 0;
    HSR_HIP(hipMemGetInfo(&free_b, &total_b));
    const bool fits = wbytes <= A.W.bytes || (double)wbytes <= 0.7 * (double)free_b;
    const int mode = c->tuning.adjoint_mode;
    A.fused = mode == 1 || (mode == 0 && fits);
    if (A.fused) HSR_HIP(dbuf_reserve(A.W, wbytes));
    HSR_HIP(dbuf_reserve(A.counts, sizeof(int64_t) * 4));
  }
  if (!A.fused) {
    k_tr_counts<<<nblk(ns, 256), 256, 0, st>>>(ns, nn, key2, val2, seg_t, cn, leaves, tcnt, start);
    if ((s = excl_scan(c, tcnt, eoff, ns)) != HSR_OK) return s;
  }
  // number of particle segments = first position whose key >= nn
  std::vector<int32_t> hkey(ns);
  HSR_HIP(hipMemcpyAsync(hkey.data(), key2, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, st));
  HSR_HIP(hipStreamSynchronize(st));
  const int64_t nsp = std::lower_bound(hkey.begin(), hkey.end(), (int32_t)nn) - hkey.begin();
  const int64_t npe = std::lower_bound(hkey.begin(), hkey.end(), (int32_t)(2 * nn)) - hkey.begin();
  A.nsp = nsp;
  A.npseudo = npe - nsp;
  const int64_t ntj = (nsp && !A.fused) ? tail_sum(c, eoff, tcnt, nsp) : 0;
  // runs (source leaves) among the particle segments
  std::vector<int64_t> hruns;
  for (int64_t p = 0; p < nsp; ++p)
    if (p == 0 || hkey[p - 1] != hkey[p]) hruns.push_back(p);
  const int64_t nruns = (int64_t)hruns.size();
  HSR_HIP(hipMemcpyAsync(run_first, hruns.data(), sizeof(int64_t) * nruns, hipMemcpyHostToDevice, st));
  A.nruns = nruns;
  A.run_first = run_first;
  if (!A.fused) {
    HSR_HIP(dbuf_reserve(A.tj, sizeof(int32_t) * (ntj + 1)));
    if (nsp) k_tr_expand<<<(unsigned)nsp, 64, 0, st>>>(nsp, val2, seg_t, cn, leaves, eoff, dptr<int32_t>(A.tj));
  }
  // source-pass items (two-pass schedule)
  HSR_HIP(dbuf_reserve(A.tr, sizeof(int64_t) * 5 * (nruns + 2)));
  int64_t *tnc = dptr<int64_t>(A.tr), *tib = tnc + (nruns + 2), *tps = tib + (nruns + 2);
  int64_t *tpb = tps + (nruns + 2), *tfirst = tpb + (nruns + 2);
  if (nruns && !A.fused) {
    k_tr_items<<<nblk(nruns, 256), 256, 0, st>>>(nruns, nsp, run_first, key2, eoff, tcnt, cn, chunk, tnc, tps);
    if ((s = excl_scan(c, tnc, tib, nruns)) != HSR_OK) return s;
    if ((s = excl_scan(c, tps, tpb, nruns)) != HSR_OK) return s;
    const int64_t acc_i = tail_sum(c, tib, tnc, nruns);
    const int64_t acc_p = tail_sum(c, tpb, tps, nruns);
    A.n_tr_items = acc_i;
    A.tr_part = acc_p;
    HSR_HIP(dbuf_reserve(A.tr_items, sizeof(AItem) * (acc_i + 1)));
    k_tr_items_fill<<<nblk(nruns, 256), 256, 0, st>>>(nruns, nsp, run_first, key2, eoff, tcnt, cn, chunk, tnc,
                                                      tib, tpb, dptr<AItem>(A.tr_items), tfirst);
    tail_items[0] = acc_i;
    HSR_HIP(hipMemcpyAsync(tfirst + nruns, &tail_items[0], sizeof(int64_t), hipMemcpyHostToDevice, st));
  } else {
    A.n_tr_items = 0;
    A.tr_part = 0;
  }
  A.tr_first = tfirst;
  A.n_tr_leaves = nruns;
  tail_items[1] = A.n_fw_items;
  HSR_HIP(hipMemcpyAsync(ffirst + nl, &tail_items[1], sizeof(int64_t), hipMemcpyHostToDevice, st));
  // canonical parents (for the pseudo push-down)
  HSR_HIP(dbuf_reserve(A.par, sizeof(int32_t) * (nn + 1) + sizeof(int64_t) * (T.nboxes + 1)));
  int32_t *par = dptr<int32_t>(A.par);
  int64_t *cb = reinterpret_cast<int64_t *>(par + nn + 1);
  HSR_HIP(hipMemcpyAsync(cb, T.box_root.data(), sizeof(int64_t) * (T.nboxes + 1), hipMemcpyHostToDevice, st));
  k_cparent<<<nblk(nn, 256), 256, 0, st>>>(nn, cn, par);
  k_roots<<<nblk(T.nboxes, 64), 64, 0, st>>>(T.nboxes, cb, par);
  tail_items[2] = A.n_fw_items;
  tail_items[3] = A.n_tr_items;
  HSR_HIP(hipMemcpyAsync(dptr<int64_t>(A.counts), &tail_items[2], 2 * sizeof(int64_t), hipMemcpyHostToDevice, st));
  HSR_HIP(hipStreamSynchronize(st));
  HSR_HIP(hipGetLastError());
  // buffers of the passes (reserved here so that a captured graph allocates nothing)
  HSR_HIP(dbuf_reserve(A.qctr, 64));
  HSR_HIP(dbuf_reserve(A.part, sizeof(float) * (std::max(A.fw_part, A.tr_part) + 1)));
  HSR_HIP(dbuf_reserve(A.H, sizeof(float) * 3 * (nn + 1) + sizeof(float) * 3 * (A.npseudo + 1)));
  A.gen = c->lists_gen;
  A.planned_mode = c->tuning.adjoint_mode;
  return HSR_OK;
}

static hsr_status adjoint_enqueue(hsr_ctx *c, const float *lx, const float *ly,
                                  const float *lz, float *gx, float *gy, float *gz);

// The adjoint passes.  The plan (transposition, item queues; host syncs) is
// made once per list build; with tuning.use_graph the passes themselves are
// captured once per plan and argument set and replayed as one graph.
hsr_status adjoint_run(hsr_ctx *c, const float *lx, const float *ly,
                       const float *lz, float *gx, float *gy, float *gz) {
  hipStream_t st = c->stream;
  AdjState &A = c->adj;
  if (c->law.guard) {
    hsr_set_error("hsr_adjoint: needs rsm > 0");
    return HSR_ERR_UNSUPPORTED;
  }
  if (c->law.kind != LAWK_POLY) {
    hsr_set_error("hsr_adjoint: implements the polynomial laws (STEP16, SR_POLY5, SR_POLY6)");
    return HSR_ERR_UNSUPPORTED;
  }
  if (A.gen != c->lists_gen || A.planned_mode != c->tuning.adjoint_mode) {
    hsr_status s = adjoint_plan(c);
    if (s != HSR_OK) return s;
    A.graph_key[0] = nullptr;  // re-capture
  }
  if (!c->tuning.use_graph) {
    HSR_HIP(hipEventRecord(c->ev[6], st));
    hsr_status s = adjoint_enqueue(c, lx, ly, lz, gx, gy, gz);
    if (s != HSR_OK) return s;
    HSR_HIP(hipEventRecord(c->ev[7], st));
    c->kernel_timed = true;
    return HSR_OK;
  }
  const void *key[6] = {lx, ly, lz, gx, gy, gz};
  bool stale = A.graph_exec == nullptr || A.graph_cull != c->tuning.cull_off ||
               A.graph_wpc != c->tuning.waves_per_cu;
  for (int k = 0; k < 6; ++k) stale = stale || A.graph_key[k] != key[k];
  if (stale) {
    if (A.graph_exec) {
      HSR_HIP(hipGraphExecDestroy(A.graph_exec));
      A.graph_exec = nullptr;
    }
    if (A.graph) {
      HSR_HIP(hipGraphDestroy(A.graph));
      A.graph = nullptr;
    }
    HSR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hsr_status s = adjoint_enqueue(c, lx, ly, lz, gx, gy, gz);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (s != HSR_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    HSR_HIP(e);
    A.graph = g;
    HSR_HIP(hipGraphInstantiate(&A.graph_exec, g, nullptr, nullptr, 0));
    for (int k = 0; k < 6; ++k) A.graph_key[k] = key[k];
    A.graph_cull = c->tuning.cull_off;
    A.graph_wpc = c->tuning.waves_per_cu;
  }
  HSR_HIP(hipEventRecord(c->ev[6], st));
  HSR_HIP(hipGraphLaunch(A.graph_exec, st));
  HSR_HIP(hipEventRecord(c->ev[7], st));
  c->kernel_timed = true;
  return HSR_OK;
}

static hsr_status adjoint_enqueue(hsr_ctx *c, const float *lx, const float *ly,
                                  const float *lz, float *gx, float *gy, float *gz) {
  hipStream_t st = c->stream;
  AdjState &A = c->adj;
  const int64_t n = c->tree.n, nn = c->tree.nnodes;
  const hsr_node *cn = dptr<hsr_node>(c->tree.canon_nodes);
  const PJ *pj = dptr<PJ>(c->tree.pj);
  AdjLaw L = adj_law(c->law);
  L.cull = c->tuning.cull_off ? 0 : 1;
  unsigned long long *q = dptr<unsigned long long>(A.qctr);
  float *part = dptr<float>(A.part);
  const int32_t *iperm = dptr<int32_t>(c->lists.iperm);
  float *H = dptr<float>(A.H);
  float *gseg = H + 3 * (nn + 1);
  const int64_t *counts = dptr<int64_t>(A.counts);
  const int wpc = c->tuning.waves_per_cu > 0 ? c->tuning.waves_per_cu : 32;
  if (n) {
    k_fill<<<nblk(n, 256), 256, 0, st>>>(n, gx);
    k_fill<<<nblk(n, 256), 256, 0, st>>>(n, gy);
    k_fill<<<nblk(n, 256), 256, 0, st>>>(n, gz);
  }
  // target pass
  if (A.n_fw_items) {
    k_zero_u64a<<<1, 64, 0, st>>>(q);
    const int64_t grid = std::min<int64_t>(A.n_fw_items, (int64_t)std::max(1, c->num_cus) * wpc);
#define LA(MODE, ORD, ITEMS, CNT)                                                          \
  k_adj<MODE, ORD><<<(unsigned)grid, 64, 0, st>>>(cn, ITEMS, CNT, MODE == 0 ?               \
      dptr<int32_t>(c->lists.jidx) : dptr<int32_t>(A.tj), pj, lx, ly, lz, L, q, gx, gy, gz, part)
    if (A.fused) {
      float *W = dptr<float>(A.W);
      const int64_t ne = c->lists.nentries + 1;
      if (L.order == 6)
        k_adj_fused6<<<(unsigned)grid, 64, 0, st>>>(cn, dptr<AItem>(A.fw_items), counts,
            dptr<int32_t>(c->lists.jidx), pj, iperm, lx, ly, lz, L, q, gx, gy, gz, part, W, W + ne,
            W + 2 * ne);
      else
        k_adj_fused5<<<(unsigned)grid, 64, 0, st>>>(cn, dptr<AItem>(A.fw_items), counts,
            dptr<int32_t>(c->lists.jidx), pj, iperm, lx, ly, lz, L, q, gx, gy, gz, part, W, W + ne,
            W + 2 * ne);
    } else if (L.order == 6) {
      LA(0, 6, dptr<AItem>(A.fw_items), counts);
    } else {
      LA(0, 5, dptr<AItem>(A.fw_items), counts);
    }
    if (A.fw_part)
      k_adj_reduce<<<(unsigned)A.n_fw_leaves, 64, 0, st>>>(cn, dptr<AItem>(A.fw_items), A.fw_first,
                                                            A.n_fw_leaves, 0, A.fused ? iperm : nullptr,
                                                            part, gx, gy, gz);
  }
  const int64_t ns = c->lists.nsegs;
  const int64_t *seg_e = dptr<int64_t>(c->lists.seg) + 2 * (ns + 1);
  const int64_t *ent_ptr = dptr<int64_t>(c->lists.ent_ptr);
  const float *W = A.fused ? dptr<float>(A.W) : nullptr;
  const int64_t ne = c->lists.nentries + 1;
  // source side of the fused schedule
  if (A.fused && A.nruns)
    k_src_reduce<<<(unsigned)A.nruns, 64, 0, st>>>(A.nruns, A.nsp, A.run_first, A.key, A.val, A.seg_t, ent_ptr,
                                                  seg_e, cn, pj, W, W + ne, W + 2 * ne, gx, gy, gz);
  // source pass of the two-pass schedule
  if (!A.fused && A.n_tr_items) {
    k_zero_u64a<<<1, 64, 0, st>>>(q);
    const int64_t grid = std::min<int64_t>(A.n_tr_items, (int64_t)std::max(1, c->num_cus) * wpc);
    if (L.order == 6) LA(1, 6, dptr<AItem>(A.tr_items), counts + 1);
    else LA(1, 5, dptr<AItem>(A.tr_items), counts + 1);
#undef LA
    if (A.tr_part)
      k_adj_reduce<<<(unsigned)A.n_tr_leaves, 64, 0, st>>>(cn, dptr<AItem>(A.tr_items), A.tr_first,
                                                            A.n_tr_leaves, 1, nullptr, part, gx, gy, gz);
  }
  // pseudo pass
  if (nn) k_fill<<<nblk(3 * nn, 256), 256, 0, st>>>(3 * nn, H);
  if (A.npseudo && A.fused) {
    k_pseudo_from_w<<<nblk(A.npseudo, 256), 256, 0, st>>>(A.npseudo, A.nsp, A.val, A.seg_t, ent_ptr, seg_e,
                                                          W, W + ne, W + 2 * ne, gseg);
    k_pseudo_node<<<(unsigned)A.npseudo, 64, 0, st>>>(A.npseudo, A.nsp, nn, A.key, gseg, cn, pj, 1, H);
    k_pseudo_push<<<(unsigned)nn, 64, 0, st>>>(nn, cn, dptr<int32_t>(A.par), H, pj, gx, gy, gz);
  } else if (A.npseudo) {
    if (L.order == 6)
      k_pseudo_grad<6><<<(unsigned)A.npseudo, 64, 0, st>>>(A.npseudo, A.nsp, nn, n, A.key, A.val, A.seg_t,
                                                           dptr<int32_t>(c->lists.leaves), cn, pj, lx, ly, lz, L, gseg);
    else
      k_pseudo_grad<5><<<(unsigned)A.npseudo, 64, 0, st>>>(A.npseudo, A.nsp, nn, n, A.key, A.val, A.seg_t,
                                                           dptr<int32_t>(c->lists.leaves), cn, pj, lx, ly, lz, L, gseg);
    k_pseudo_node<<<(unsigned)A.npseudo, 64, 0, st>>>(A.npseudo, A.nsp, nn, A.key, gseg, cn, pj, 0, H);
    k_pseudo_push<<<(unsigned)nn, 64, 0, st>>>(nn, cn, dptr<int32_t>(A.par), H, pj, gx, gy, gz);
  }
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}
