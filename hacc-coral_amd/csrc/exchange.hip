// exchange.hip -- the overload refresh of one rank (MC3Extras::particleExtras
// refresh branch, MC3Extras.cxx:330-338, 660-700) with the particles on the
// device:
//
//   to physical units and keep the alive ones   Particles::copyAliveIntoVectors
//                                               (Particles.cxx:798-830, 945-1030)
//   copy the border particles for the 26        ParticleExchange::exchangeParticles
//   neighbours, shifted by the periodic offsets (ParticleExchange.cxx:488-728)
//   unit masses, back to grid units             Particles::copyAllFromVectors
//                                               (Particles.cxx:298-366, 762-796)
//   cell sort                                   Particles::resortParticles
//
// With one rank every neighbour is the rank itself, so the exchange appends
// the copies to its own arrays (ParticleExchange.cxx:676-695): the alive
// particles in input order, then neighbour 0's copies, neighbour 1's, ...
// each in particle order.  The output position of every particle is thus an
// exclusive prefix sum over (category, particle) with 27 categories (alive,
// 26 neighbours) in category-major order.  Each wave owns a tile of 1024
// consecutive particles; one pass counts the tile's members of each category,
// a device scan turns the 27 x tiles counts into output offsets, and a second
// pass recomputes the same classification and writes every record at its
// offset plus its rank among the wave's members (ballot prefix).  The
// classification is pure float arithmetic on the inputs, so both passes
// agree.  Every pass streams: HBM-bound.
//
// Built without FMA contraction: (x + corner) * grid2phys_pos and
// x * phys2grid_pos - corner round after each operation as on x86.
#include "hsr_internal.h"

#include <hipcub/hipcub.hpp>

#include <cmath>

namespace {

constexpr int XITEMS = 16;           // particles per lane per tile
constexpr int XTILE = 64 * XITEMS;   // particles per wave tile
constexpr int NCAT = 27;             // alive + 26 neighbours
constexpr int NNB = 26;

// Direction of each neighbour per dimension (-1 low face, +1 high face, 0
// not involved), in the enum order of BasicDefinition.h:224-264
// (X0, X1, Y0, Y1, Z0, Z1, X0_Y0, X1_Y1, X0_Y1, X1_Y0, Y0_Z0, Y1_Z1, Y0_Z1,
// Y1_Z0, Z0_X0, Z1_X1, Z0_X1, Z1_X0, X0_Y0_Z0, X1_Y1_Z1, X0_Y0_Z1, X1_Y1_Z0,
// X0_Y1_Z0, X1_Y0_Z1, X0_Y1_Z1, X1_Y0_Z0).
constexpr int kDir[NNB][3] = {
    {-1, 0, 0},  {1, 0, 0},  {0, -1, 0},  {0, 1, 0},   {0, 0, -1},  {0, 0, 1},
    {-1, -1, 0}, {1, 1, 0},  {-1, 1, 0},  {1, -1, 0},  {0, -1, -1}, {0, 1, 1},
    {0, -1, 1},  {0, 1, -1}, {-1, 0, -1}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},
    {-1, -1, -1}, {1, 1, 1}, {-1, -1, 1}, {1, 1, -1},  {-1, 1, -1}, {1, -1, 1},
    {-1, 1, 1},  {1, -1, -1}};

// Everything the classification and the unit conversions read, passed by
// value (kernel arguments).
struct XTable {
  float corner;               // corner_grid_total as float (-ng_overload)
  float g2p, p2g;             // grid2phys_pos, phys2grid_pos
  float vl2g, vg2l;           // velocity factors of local2global, global2local
  float lo[3], hi[3];         // alive region [lo, hi)
  float mlo[3], mhi[3];       // interior (mlo, mhi): never shared
  float rmin[NNB][3], rmax[NNB][3];  // closed region sent to each neighbour
  float off[NNB][3];          // position shift of the copies
};

__device__ __forceinline__ void to_phys(const XTable &T, float x, float y, float z,
                                        float &gx, float &gy, float &gz) {
  gx = x + T.corner;
  gx = gx * T.g2p;
  gy = y + T.corner;
  gy = gy * T.g2p;
  gz = z + T.corner;
  gz = gz * T.g2p;
}

__device__ __forceinline__ float to_grid(const XTable &T, float g) {
  float x = g * T.p2g;
  return x - T.corner;
}

// Alive test (Particles.cxx:965-972) and the neighbours whose region holds
// the particle (ParticleExchange.cxx:542-571).
__device__ __forceinline__ bool classify(const XTable &T, float gx, float gy, float gz,
                                         uint32_t &bits) {
  bits = 0;
  const bool alive = gx >= T.lo[0] && gx < T.hi[0] && gy >= T.lo[1] && gy < T.hi[1] &&
                     gz >= T.lo[2] && gz < T.hi[2];
  if (!alive) return false;
  if (gx > T.mlo[0] && gx < T.mhi[0] && gy > T.mlo[1] && gy < T.mhi[1] && gz > T.mlo[2] &&
      gz < T.mhi[2])
    return true;
#pragma unroll
  for (int n = 0; n < NNB; ++n)
    if (gx >= T.rmin[n][0] && gx <= T.rmax[n][0] && gy >= T.rmin[n][1] &&
        gy <= T.rmax[n][1] && gz >= T.rmin[n][2] && gz <= T.rmax[n][2])
      bits |= 1u << n;
  return true;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = threadIdx.x & 63;
  return l ? (~0ull >> (64 - l)) : 0ull;
}

// cnt[k * ntiles + w] = members of category k in tile w.
__global__ void __launch_bounds__(256)
k_xchg_count(int64_t n, int64_t ntiles, const float *__restrict__ x,
             const float *__restrict__ y, const float *__restrict__ z, XTable T,
             int64_t *__restrict__ cnt) {
  const int64_t w = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (w >= ntiles) return;  // whole waves
  const int l = threadIdx.x & 63;
  int c[NCAT];
#pragma unroll
  for (int k = 0; k < NCAT; ++k) c[k] = 0;
  // all of the tile's positions in flight at once (one memory round trip)
  float px[XITEMS], py[XITEMS], pz[XITEMS];
#pragma unroll
  for (int it = 0; it < XITEMS; ++it) {
    const int64_t i = w * XTILE + it * 64 + l;
    px[it] = i < n ? x[i] : 0.f;
    py[it] = i < n ? y[i] : 0.f;
    pz[it] = i < n ? z[i] : 0.f;
  }
#pragma unroll
  for (int it = 0; it < XITEMS; ++it) {
    const int64_t i = w * XTILE + it * 64 + l;
    bool al = false;
    uint32_t bits = 0;
    if (i < n) {
      float gx, gy, gz;
      to_phys(T, px[it], py[it], pz[it], gx, gy, gz);
      al = classify(T, gx, gy, gz, bits);
    }
    c[0] += __popcll(__ballot(al));
    if (__ballot(bits != 0)) {
#pragma unroll
      for (int k = 0; k < NNB; ++k) c[k + 1] += __popcll(__ballot((bits >> k) & 1u));
    }
  }
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < NCAT; ++k) cnt[k * ntiles + w] = c[k];
  }
}

// tot[0] = all records, tot[1] = alive records.
__global__ void k_xchg_total(int64_t m, int64_t ntiles, const int64_t *__restrict__ cnt,
                             const int64_t *__restrict__ off, int64_t *__restrict__ tot) {
  if (threadIdx.x == 0) {
    tot[0] = off[m - 1] + cnt[m - 1];
    tot[1] = off[ntiles];
  }
}

struct XOut {
  float *f[8];
  int64_t *id;
  uint16_t *mask;
};
struct XIn {
  const float *f[8];
  const int64_t *id;
  const uint16_t *mask;
};

__device__ __forceinline__ void put(const XOut &o, int64_t p, float x, float y, float z,
                                    float vx, float vy, float vz, float phi, int64_t id,
                                    uint16_t mask) {
  o.f[0][p] = x;
  o.f[1][p] = y;
  o.f[2][p] = z;
  o.f[3][p] = vx;
  o.f[4][p] = vy;
  o.f[5][p] = vz;
  o.f[6][p] = 1.0f;
  o.f[7][p] = phi;
  o.id[p] = id;
  o.mask[p] = mask;
}

__global__ void __launch_bounds__(256)
k_xchg_scatter(int64_t n, int64_t ntiles, XIn in, XOut out, XTable T,
               const int64_t *__restrict__ off) {
  const int64_t w = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (w >= ntiles) return;
  const int l = threadIdx.x & 63;
  const uint64_t lt = lanemask_lt();
  int64_t base[NCAT];
#pragma unroll
  for (int k = 0; k < NCAT; ++k) base[k] = off[k * ntiles + w];
#pragma unroll 2
  for (int it = 0; it < XITEMS; ++it) {
    const int64_t i = w * XTILE + it * 64 + l;
    bool al = false;
    uint32_t bits = 0;
    float gx = 0.f, gy = 0.f, gz = 0.f, vx = 0.f, vy = 0.f, vz = 0.f, phi = 0.f;
    int64_t id = 0;
    uint16_t mk = 0;
    if (i < n) {
      // every field loaded up front (almost all particles are kept), so the
      // item costs one memory round trip
      const float x0 = in.f[0][i], y0 = in.f[1][i], z0 = in.f[2][i];
      const float u = in.f[3][i], v = in.f[4][i], t = in.f[5][i];
      phi = in.f[7][i];
      id = in.id[i];
      mk = in.mask[i];
      to_phys(T, x0, y0, z0, gx, gy, gz);
      al = classify(T, gx, gy, gz, bits);
      // physical velocities, then back to grid units: two roundings each
      vx = u * T.vl2g;
      vx = vx * T.vg2l;
      vy = v * T.vl2g;
      vy = vy * T.vg2l;
      vz = t * T.vl2g;
      vz = vz * T.vg2l;
    }
    const uint64_t ba = __ballot(al);
    if (al)
      put(out, base[0] + __popcll(ba & lt), to_grid(T, gx), to_grid(T, gy), to_grid(T, gz), vx,
          vy, vz, phi, id, mk);
    base[0] += __popcll(ba);
    if (__ballot(bits != 0)) {
#pragma unroll
      for (int k = 0; k < NNB; ++k) {
        const bool m = (bits >> k) & 1u;
        const uint64_t bk = __ballot(m);
        if (m) {
          const float sx = gx + T.off[k][0], sy = gy + T.off[k][1], sz = gz + T.off[k][2];
          put(out, base[k + 1] + __popcll(bk & lt), to_grid(T, sx), to_grid(T, sy),
              to_grid(T, sz), vx, vy, vz, phi, id, mk);
        }
        base[k + 1] += __popcll(bk);
      }
    }
  }
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// The rank's tables, with the reference's float arithmetic on the host:
// Domain.cxx:65-78 (one rank: decomposition 1, position 0),
// ParticleExchange.cxx:141-229 (offsets) and 243-443 (regions),
// Particles.cxx:769-773 and 805-809 (unit factors).
XTable make_table(const hsr_domain &d, float anow) {
  XTable T{};
  const int corner_grid_alive = 0;
  const int corner_grid_total = corner_grid_alive - d.ng_overload;
  T.corner = (float)corner_grid_total;
  T.g2p = d.grid2phys_pos;
  T.p2g = d.phys2grid_pos;
  T.vl2g = d.grid2phys_vel / anow / anow;
  T.vg2l = d.phys2grid_vel * anow * anow;
  const float ca = (float)corner_grid_alive * d.grid2phys_pos;  // corner_phys_alive
  const float rla = d.rL / 1;                                    // rL_local_alive
  const float box = d.rL, dead = d.oL;
  const float step = box / 1;
  const float share_lo = 0 * step;
  float share_hi = share_lo + step;
  if (share_hi > box) share_hi = box;
  const float mine_lo = share_lo + dead;
  const float mine_hi = share_hi - dead;
  for (int k = 0; k < 3; ++k) {
    T.lo[k] = ca;
    T.hi[k] = ca + rla;
    T.mlo[k] = mine_lo;
    T.mhi[k] = mine_hi;
  }
  for (int n = 0; n < NNB; ++n)
    for (int k = 0; k < 3; ++k) {
      const int dir = kDir[n][k];
      T.rmin[n][k] = dir > 0 ? mine_hi : share_lo;
      T.rmax[n][k] = dir < 0 ? mine_lo : share_hi;
      // the low-face neighbour of the first rank gets +rL, the high-face one
      // of the last rank -rL (one rank is both)
      const int factor = dir < 0 ? 1 : (dir > 0 ? -1 : 0);
      T.off[n][k] = factor * box;
    }
  return T;
}

}  // namespace

// Domain::initialize (Domain.cxx:34-84) for rL, oL and ng; float and double
// operations as written there.
void domain_derive(hsr_domain *d) {
  const float rL = d->rL;
  const int ng = d->ng;
  d->phys2grid_pos = ng / rL;
  d->grid2phys_pos = rL / ng;
  d->phys2grid_vel = (float)(ng / (100.0 * rL));
  d->grid2phys_vel = (float)(100.0 * rL / ng);
  d->ng_overload = static_cast<int>(ceilf(d->oL_in * d->phys2grid_pos));
  d->oL = d->ng_overload * d->grid2phys_pos;
  d->ng_local_total = ng + 2 * d->ng_overload + 1;
}

hsr_status refresh_count(hsr_ctx *c, const hsr_domain *dom, float anow, int64_t n,
                         const float *x, const float *y, const float *z, int64_t *n_out,
                         int64_t *alive) {
  *n_out = 0;
  *alive = 0;
  if (n <= 0) return HSR_OK;
  hipStream_t st = c->stream;
  const XTable T = make_table(*dom, anow);
  const int64_t ntiles = (n + XTILE - 1) / XTILE;
  const int64_t m = NCAT * ntiles;
  HSR_HIP(dbuf_reserve(c->sub.xcnt, sizeof(int64_t) * (2 * (size_t)m + 2) + 256));
  int64_t *cnt = dptr<int64_t>(c->sub.xcnt);
  int64_t *off = cnt + m;
  int64_t *tot = off + m;
  k_xchg_count<<<nblk(ntiles * 64, 256), 256, 0, st>>>(n, ntiles, x, y, z, T, cnt);
  HSR_HIP(hipGetLastError());
  size_t tmp = 0;
  HSR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt, off, (int)m, st));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp));
  HSR_HIP(hipcub::DeviceScan::ExclusiveSum(c->s.cub.p, tmp, cnt, off, (int)m, st));
  k_xchg_total<<<1, 64, 0, st>>>(m, ntiles, cnt, off, tot);
  HSR_HIP(hipGetLastError());
  int64_t h[2] = {0, 0};
  HSR_HIP(hipMemcpyAsync(h, tot, sizeof(h), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipStreamSynchronize(st));
  *n_out = h[0];
  *alive = h[1];
  return HSR_OK;
}

hsr_status refresh_finish(hsr_ctx *c, const hsr_domain *dom, float anow, int64_t n,
                          const PArrays &in, int64_t n_out, const PArrays &out,
                          int64_t *inside) {
  *inside = 0;
  if (n <= 0 || n_out <= 0) return HSR_OK;
  hipStream_t st = c->stream;
  const XTable T = make_table(*dom, anow);
  const int64_t ntiles = (n + XTILE - 1) / XTILE;
  const int64_t m = NCAT * ntiles;
  const int64_t *off = dptr<int64_t>(c->sub.xcnt) + m;
  // staging: id (8 B) first keeps every array aligned, then 8 floats, mask
  const size_t no = (size_t)n_out;
  HSR_HIP(dbuf_reserve(c->sub.xstage, (8 + 8 * 4 + 2) * no + 256));
  char *sp = (char *)c->sub.xstage.p;
  PArrays mid{};
  mid.id = (int64_t *)sp;
  sp += sizeof(int64_t) * no;
  for (int a = 0; a < 8; ++a) {
    mid.f[a] = (float *)sp;
    sp += sizeof(float) * no;
  }
  mid.mask = (uint16_t *)sp;
  XIn xi{};
  XOut xo{};
  for (int a = 0; a < 8; ++a) {
    xi.f[a] = in.f[a];
    xo.f[a] = mid.f[a];
  }
  xi.id = in.id;
  xi.mask = in.mask;
  xo.id = mid.id;
  xo.mask = mid.mask;
  k_xchg_scatter<<<nblk(ntiles * 64, 256), 256, 0, st>>>(n, ntiles, xi, xo, T, off);
  HSR_HIP(hipGetLastError());
  const int32_t ng[3] = {dom->ng_local_total, dom->ng_local_total, dom->ng_local_total};
  return resort_into(c, n_out, ng, mid, out, inside);
}
