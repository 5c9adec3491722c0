// pm.hip -- the particle-grid steps either side of the PM solve, on the
// device: the cloud-in-cell deposit (Particles::cic, Particles.cxx:589-641)
// and its interpolation back to the particles (Particles::inverse_cic,
// Particles.cxx:645-714).
//
// Both follow the reference's arithmetic exactly: the weights ab, de, gh
// are 1.0 + (ix - x) with a float difference and a double sum stored as
// float; of the eight corner terms the first is a float product and the
// other seven are double from their first (1.0 - w) factor on, and each is
// added to its float accumulator in that precision.  That is the
// reference's own use of double (its literals 1.0 promote the arithmetic);
// compiled without FMA contraction.
//
// The deposit is a scatter whose result depends on the order of the float
// additions into each cell.  To reproduce the reference's sequential order
// in parallel it sorts the particles stably by base cell; a grid cell's
// contributions are then the merge, by particle index, of the particle
// lists of its eight neighbouring base cells.  A cell with at most 64 is
// merged and added by its own thread; a longer cell (dense clumps) by a
// workgroup, which forms its terms in merged order into LDS and adds them,
// 256 at a time, through exact fixed-grid maps of the float accumulation;
// the slot for corners outside the grid gets its contributions in
// (particle, corner) order and a workgroup.  The interpolation is a gather, one thread per
// particle.
#include "hsr_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

namespace {

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// (int)floorf(x) as the reference's x86 build converts it: NaN and values
// outside the int range become INT_MIN (the conversion's indefinite value).
__device__ __forceinline__ int x86_floor_int(float x) {
  const float f = floorf(x);
  return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
}

struct Grid {
  int ng0, ng1, ng2;
  uint32_t safe;  // ng0 * ng1 * ng2: the slot of corners outside the grid
};

// Particles.cxx:371-395 with lo = 0, hi = ng.
__device__ __forceinline__ uint32_t cell_of(int ix, int iy, int iz, const Grid &g) {
  if (ix >= 0 && ix < g.ng0 && iy >= 0 && iy < g.ng1 && iz >= 0 && iz < g.ng2)
    return ((uint32_t)ix * (uint32_t)g.ng1 + (uint32_t)iy) * (uint32_t)g.ng2 + (uint32_t)iz;
  return g.safe;
}

struct Corners {
  uint32_t cell[8];
  float ab, de, gh;
};

// Particles.cxx:617-640: corners in the reference's order.
__device__ __forceinline__ Corners corners(float xx, float yy, float zz, const Grid &g) {
  Corners c;
  const int ix = x86_floor_int(xx), iy = x86_floor_int(yy), iz = x86_floor_int(zz);
  const int ip = ix + 1, jp = iy + 1, kp = iz + 1;
  c.ab = (float)(1.0 + (double)((float)ix - xx));
  c.de = (float)(1.0 + (double)((float)iy - yy));
  c.gh = (float)(1.0 + (double)((float)iz - zz));
  c.cell[0] = cell_of(ix, iy, iz, g);
  c.cell[1] = cell_of(ix, jp, iz, g);
  c.cell[2] = cell_of(ix, jp, kp, g);
  c.cell[3] = cell_of(ix, iy, kp, g);
  c.cell[4] = cell_of(ip, iy, kp, g);
  c.cell[5] = cell_of(ip, jp, kp, g);
  c.cell[6] = cell_of(ip, jp, iz, g);
  c.cell[7] = cell_of(ip, iy, iz, g);
  return c;
}

// Corner q's term with lead factor s, evaluated as the reference's
// expression is: float products until the first (1.0 - w), double after.
__device__ __forceinline__ double corner_term(int q, float s, float ab, float de, float gh) {
  switch (q) {
  case 0: return (double)(s * ab * de * gh);
  case 1: return (double)(s * ab) * (1.0 - (double)de) * (double)gh;
  case 2: return (double)(s * ab) * (1.0 - (double)de) * (1.0 - (double)gh);
  case 3: return (double)(s * ab * de) * (1.0 - (double)gh);
  case 4: return (double)s * (1.0 - (double)ab) * (double)de * (1.0 - (double)gh);
  case 5: return (double)s * (1.0 - (double)ab) * (1.0 - (double)de) * (1.0 - (double)gh);
  case 6: return (double)s * (1.0 - (double)ab) * (1.0 - (double)de) * (double)gh;
  default: return (double)s * (1.0 - (double)ab) * (double)de * (double)gh;
  }
}

// The reference's accumulation step for one contribution
// (Particles.cxx:632-639): corner 0 adds a float term in float, the others
// add a double term and round to float.  The float addition r + t equals
// (float)((double)r + t) for float r and t (double rounding through a
// format of at least 2 * 24 + 2 bits is innocuous for a sum), so every
// corner takes the one form below.
__device__ __forceinline__ float cic_add(float r, double t) {
  return (float)((double)r + t);
}

__device__ __forceinline__ double readlane_f64(double v, int i) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), i);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), i);
  return __hiloint2double(hi, lo);
}

// Fixed-grid maps of the accumulation r <- (float)((double)r + t), t >= 0.
// While r stays in one binade [2^E, 2^(E+1)) (E >= -125) it is R u with
// u = 2^(E-23) and R an integer below 2^24, and one step is R <- R + D where
// D depends on t and on the parity of R alone: the double sum rounds to the
// grid u 2^-29 (its ulp in that binade, ties to even independent of R), the
// float to the grid u with ties to the even significand.  Maps (d0, d1) =
// (D for even R, D for odd R) compose associatively, so a wave scans the
// maps of 64 terms and applies the longest prefix that stays in the binade;
// the term that leaves it is added directly.  Totals past 2^26 saturate
// (they leave the binade anyway).
struct PMap {
  int d0, d1;
};
constexpr int PMAP_SAT = 1 << 26;

__device__ __forceinline__ int pmap_add(int a, int b) {
  const long long s = (long long)a + b;
  return s > PMAP_SAT ? PMAP_SAT : (int)s;
}

// a, then b
__device__ __forceinline__ PMap pmap_compose(PMap a, PMap b) {
  PMap m;
  m.d0 = a.d0 >= PMAP_SAT ? PMAP_SAT : pmap_add(a.d0, (a.d0 & 1) ? b.d1 : b.d0);
  m.d1 = a.d1 >= PMAP_SAT ? PMAP_SAT : pmap_add(a.d1, ((a.d1 + 1) & 1) ? b.d1 : b.d0);
  return m;
}

__device__ __forceinline__ PMap pmap_term(double t, int E) {
  PMap m;
  const double q = ldexp(t, 23 - E);                 // t / u, exact
  if (!(q < 16777216.0)) {
    m.d0 = m.d1 = PMAP_SAT;
    return m;
  }
  const double qq = ldexp(rint(ldexp(q, 29)), -29);  // the double sum's rounding
  const double a = floor(qq), f = qq - a;
  const int ai = (int)a;
  const int up = f > 0.5 ? 1 : 0;
  const bool tie = f == 0.5;
  m.d0 = ai + up + ((tie && (ai & 1)) ? 1 : 0);
  m.d1 = ai + up + ((tie && !(ai & 1)) ? 1 : 0);
  return m;
}

// One wave adds 64 LT terms of a cell's contributions in order: lane l
// holds terms LT l .. LT l + LT - 1, the first cnt of the step live.  With
// all terms nonnegative (the deposit's weights are) it applies them by the
// map scan above -- each lane composes its LT terms in order, then the
// lanes' maps are scanned -- and applies the longest prefix that stays in
// the binade; in the lane where the sum leaves it the terms are stepped one
// at a time to the crossing term, which is added directly.  Otherwise the
// terms are added in order, reading them with v_readlane.  r is the same in
// every lane.
constexpr int LT = 4;

__device__ __forceinline__ double pick(const double (&t)[LT], int k) {
  return k == 0 ? t[0] : k == 1 ? t[1] : k == 2 ? t[2] : t[3];
}

__device__ __forceinline__ float wave_add_step(float r, const double (&t)[LT], int cnt) {
  static_assert(LT == 4, "pick() reads four terms per lane");
  const int l = threadIdx.x & 63;
  bool neg = false;
#pragma unroll
  for (int k = 0; k < LT; ++k) {
    const int i = LT * l + k;
    neg = neg || (i < cnt && !(t[k] >= 0.0));
  }
  if (__ballot(neg) != 0ull) {
    for (int i = 0; i < cnt; ++i) r = cic_add(r, readlane_f64(pick(t, i % LT), i / LT));
    return r;
  }
  int s = 0;
  while (s < cnt) {
    if (!(r >= 0x1p-125f)) {  // zero or subnormal: the double sum's grid varies
      r = cic_add(r, readlane_f64(pick(t, s % LT), s / LT));
      ++s;
      continue;
    }
    const int E = ilogbf(r);
    const double u = ldexp(1.0, E - 23);
    const int R = (int)((double)r / u);
    PMap mp = PMap{0, 0};
#pragma unroll
    for (int k = 0; k < LT; ++k) {
      const int i = LT * l + k;
      if (i >= s && i < cnt) mp = pmap_compose(mp, pmap_term(t[k], E));
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      PMap pv;
      pv.d0 = __shfl_up(mp.d0, off);
      pv.d1 = __shfl_up(mp.d1, off);
      if (l >= off) mp = pmap_compose(pv, mp);
    }
    const int d = (R & 1) ? mp.d1 : mp.d0;
    const int Rl = d >= PMAP_SAT ? PMAP_SAT : R + d;
    // R only grows: the lanes past the crossing one cross too
    const unsigned long long cross = __ballot(Rl >= (1 << 24));
    if (!cross) {
      r = (float)((double)__shfl(Rl, 63) * u);
      s = cnt;
      continue;
    }
    const int j = __ffsll((long long)cross) - 1;
    // the lanes before j applied, then lane j's terms one at a time
    int Rj = j > 0 ? __shfl(Rl, j - 1) : R;
    for (int k = 0; k < LT; ++k) {  // uniform
      const int i = LT * j + k;
      if (i < s || i >= cnt) continue;
      const double tk = readlane_f64(pick(t, k), j);
      const PMap m1 = pmap_term(tk, E);
      const int dk = (Rj & 1) ? m1.d1 : m1.d0;
      if (dk >= PMAP_SAT || Rj + dk >= (1 << 24)) {  // the crossing term
        r = cic_add((float)((double)Rj * u), tk);
        s = i + 1;
        break;
      }
      Rj += dk;
    }
  }
  return r;
}

constexpr int TAIL_THREADS = 1024;

// The outside slot's terms [0, m) by one workgroup: LT TAIL_THREADS terms per step, LT
// per thread composed in order, maps scanned within each wave and across
// the 16 waves through LDS, the longest prefix that stays in the binade
// applied; the thread holding the crossing term steps its terms to it and
// adds it directly.
__global__ void __launch_bounds__(TAIL_THREADS)
k_cic_tail(int64_t m, const uint32_t *__restrict__ key, const double *__restrict__ ts,
           float *__restrict__ rho) {
  __shared__ PMap wtot[TAIL_THREADS / 64];
  __shared__ float r_sh;
  __shared__ int j_sh, neg_sh, s_sh, rj_sh;
  const int64_t p0 = 0;
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  if (tid == 0) r_sh = 0.f;
  __syncthreads();
  for (int64_t base = p0; base < m; base += LT * TAIL_THREADS) {
    const int cnt = (int)min<int64_t>(LT * TAIL_THREADS, m - base);
    double t[LT];
    bool neg = false;
#pragma unroll
    for (int k = 0; k < LT; ++k) {
      const int i = LT * tid + k;
      t[k] = i < cnt ? ts[base + i] : 0.0;
      neg = neg || (i < cnt && !(t[k] >= 0.0));
    }
    if (tid == 0) neg_sh = 0;
    __syncthreads();
    if (neg) neg_sh = 1;
    __syncthreads();
    if (neg_sh) {
      if (tid == 0) {
        float r = r_sh;
        for (int i = 0; i < cnt; ++i) r = cic_add(r, ts[base + i]);
        r_sh = r;
      }
      __syncthreads();
      continue;
    }
    int s = 0;
    while (s < cnt) {
      const float r = r_sh;
      if (!(r >= 0x1p-125f)) {
        __syncthreads();
        if (tid == 0) r_sh = cic_add(r, ts[base + s]);
        __syncthreads();
        ++s;
        continue;
      }
      const int E = ilogbf(r);
      const double u = ldexp(1.0, E - 23);
      const int R = (int)((double)r / u);
      PMap mp = PMap{0, 0};
#pragma unroll
      for (int k = 0; k < LT; ++k) {
        const int i = LT * tid + k;
        if (i >= s && i < cnt) mp = pmap_compose(mp, pmap_term(t[k], E));
      }
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        PMap pv;
        pv.d0 = __shfl_up(mp.d0, off);
        pv.d1 = __shfl_up(mp.d1, off);
        if (l >= off) mp = pmap_compose(pv, mp);
      }
      if (l == 63) wtot[w] = mp;
      if (tid == 0) j_sh = TAIL_THREADS;
      __syncthreads();
      if (w == 0) {  // exclusive prefix of the wave totals, in wave order
        PMap tv = l < TAIL_THREADS / 64 ? wtot[l] : PMap{0, 0};
#pragma unroll
        for (int off = 1; off < TAIL_THREADS / 64; off <<= 1) {
          PMap pv;
          pv.d0 = __shfl_up(tv.d0, off);
          pv.d1 = __shfl_up(tv.d1, off);
          if (l >= off) tv = pmap_compose(pv, tv);
        }
        PMap ex;
        ex.d0 = __shfl_up(tv.d0, 1);
        ex.d1 = __shfl_up(tv.d1, 1);
        if (l == 0) ex = PMap{0, 0};
        if (l < TAIL_THREADS / 64) wtot[l] = ex;
      }
      __syncthreads();
      const PMap ex = wtot[w];  // the waves before this one
      const PMap inc = pmap_compose(ex, mp);
      const int d = (R & 1) ? inc.d1 : inc.d0;
      const int Rl = d >= PMAP_SAT ? PMAP_SAT : R + d;
      // R only grows: the threads past the crossing one cross too
      if (Rl >= (1 << 24)) atomicMin(&j_sh, tid);
      __syncthreads();
      const int j = j_sh;
      if (j == TAIL_THREADS) {  // no crossing: all applied
        if (tid == TAIL_THREADS - 1) r_sh = (float)((double)Rl * u);
        s = cnt;
        __syncthreads();
        continue;
      }
      // the threads before j applied: R after them is thread j - 1's
      // inclusive value; then thread j steps its own terms to the crossing
      if (tid == j - 1) rj_sh = Rl;
      if (j == 0 && tid == 0) rj_sh = R;
      __syncthreads();
      if (tid == j) {
        int Rj = rj_sh;
        for (int k = 0; k < LT; ++k) {
          const int i = LT * tid + k;
          if (i < s || i >= cnt) continue;
          const PMap m1 = pmap_term(t[k], E);
          const int dk = (Rj & 1) ? m1.d1 : m1.d0;
          if (dk >= PMAP_SAT || Rj + dk >= (1 << 24)) {  // the crossing term
            r_sh = cic_add((float)((double)Rj * u), t[k]);
            s_sh = i + 1;
            break;
          }
          Rj += dk;
        }
      }
      __syncthreads();
      s = s_sh;
      __syncthreads();
    }
  }
  if (tid == 0) rho[key[p0]] = r_sh;
}

// ------------------------------------------------- deposit by base cells
// A particle with floor coordinates (ix, iy, iz) deposits into the cells
// (ix, iy, iz) + d_q for its eight corners q (corners()), so a grid cell c
// receives, from each q, the particles whose base cell is c - d_q.  With
// the particles sorted stably by base cell, each such list is in particle
// order, and c's contributions in the reference's order (particle by
// particle; one particle reaches c through one corner at most) are the
// merge of its eight lists by particle index.  Base cells run over
// [-1, ng - 1] per axis (an extended grid, shifted by one); a particle
// outside it has all its corners outside the grid.  Corners outside the
// grid go to the outside slot, in (particle, corner) order.

// corner offsets d_q in corners()' order
__constant__ int kCornerDx[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__constant__ int kCornerDy[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int kCornerDz[8] = {0, 0, 1, 1, 1, 1, 0, 0};

struct Ext {
  uint32_t e1, e2;  // extended sizes ng1 + 1, ng2 + 1 (ng0 + 1 along x)
  uint32_t out;     // (ng0 + 1)(ng1 + 1)(ng2 + 1): the key of particles outside
};

// The weights of corners() alone.
__device__ __forceinline__ void weights(float xx, float yy, float zz, float &ab, float &de,
                                        float &gh) {
  const int ix = x86_floor_int(xx), iy = x86_floor_int(yy), iz = x86_floor_int(zz);
  ab = (float)(1.0 + (double)((float)ix - xx));
  de = (float)(1.0 + (double)((float)iy - yy));
  gh = (float)(1.0 + (double)((float)iz - zz));
}

// Per particle: its extended base-cell key, its index (the sort's value) and
// the number of its corners outside the grid.
__global__ void __launch_bounds__(256)
k_cic_base(int64_t n, const float *__restrict__ x, const float *__restrict__ y,
           const float *__restrict__ z, Grid g, Ext ex, uint32_t *__restrict__ key,
           int32_t *__restrict__ idx, int32_t *__restrict__ nout) {
  const int64_t nn = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (nn >= n) return;
  const float xx = x[nn], yy = y[nn], zz = z[nn];
  const Corners cc = corners(xx, yy, zz, g);
  int no = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) no += cc.cell[q] == g.safe;
  const int ix = x86_floor_int(xx), iy = x86_floor_int(yy), iz = x86_floor_int(zz);
  const bool in = ix >= -1 && ix < g.ng0 && iy >= -1 && iy < g.ng1 && iz >= -1 && iz < g.ng2;
  key[nn] = in ? (((uint32_t)(ix + 1) * ex.e1 + (uint32_t)(iy + 1)) * ex.e2 + (uint32_t)(iz + 1)) : ex.out;
  idx[nn] = (int32_t)nn;
  nout[nn] = no;
}

// [beg, end) of every base cell in the sorted order (left 0 when empty).
__global__ void __launch_bounds__(256)
k_cic_ranges(int64_t n, const uint32_t *__restrict__ key, Ext ex, int32_t *__restrict__ beg,
             int32_t *__restrict__ end) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t k = key[p];
  if (k == ex.out) return;
  if (p == 0 || key[p - 1] != k) beg[k] = (int32_t)p;
  if (p == n - 1 || key[p + 1] != k) end[k] = (int32_t)(p + 1);
}

constexpr int CELL_SHORT = 64;  // contributions a cell's own thread merges and adds

// The eight source lists of grid cell c = (i, j, k).
__device__ __forceinline__ void cell_lists(int i, int j, int k, const Ext &ex,
                                           const int32_t *beg, const int32_t *end,
                                           int32_t (&s)[8], int32_t (&e)[8]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t b = ((uint32_t)(i - kCornerDx[q] + 1) * ex.e1 + (uint32_t)(j - kCornerDy[q] + 1)) * ex.e2 +
                       (uint32_t)(k - kCornerDz[q] + 1);
    s[q] = beg[b];
    e[q] = end[b];
  }
}

// The next contribution of the merge: the list whose head particle comes
// first (lists never share a particle).
__device__ __forceinline__ int merge_pick(const int32_t (&hv)[8]) {
  int qs = 0;
  int32_t best = hv[0];
#pragma unroll
  for (int q = 1; q < 8; ++q)
    if (hv[q] < best) {
      best = hv[q];
      qs = q;
    }
  return qs;
}

__device__ __forceinline__ void merge_advance(int qs, int32_t (&s)[8], const int32_t (&e)[8],
                                              int32_t (&hv)[8], const int32_t *idx) {
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q == qs) {
      ++s[q];
      hv[q] = s[q] < e[q] ? idx[s[q]] : INT_MAX;
    }
}

// One thread per grid cell: at most CELL_SHORT contributions are merged and
// added here; longer cells are queued with their counts for k_cic_long_sum.
__global__ void __launch_bounds__(256)
k_cic_cells(Grid g, Ext ex, const int32_t *__restrict__ beg, const int32_t *__restrict__ end,
            const int32_t *__restrict__ idx, const float *__restrict__ x,
            const float *__restrict__ y, const float *__restrict__ z, float c,
            float *__restrict__ rho, uint32_t *__restrict__ long_cell,
            int64_t *__restrict__ long_k, unsigned long long *__restrict__ nlong) {
  const int64_t cid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (cid >= (int64_t)g.safe) return;
  const int k = (int)(cid % g.ng2);
  const int64_t ij = cid / g.ng2;
  const int j = (int)(ij % g.ng1), i = (int)(ij / g.ng1);
  int32_t s[8], e[8];
  cell_lists(i, j, k, ex, beg, end, s, e);
  int K = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) K += e[q] - s[q];
  if (K == 0) return;
  if (K > CELL_SHORT) {
    const unsigned long long slot = atomicAdd(nlong, 1ull);
    long_cell[slot] = (uint32_t)cid;
    long_k[slot] = K;
    return;
  }
  int32_t hv[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) hv[q] = s[q] < e[q] ? idx[s[q]] : INT_MAX;
  float r = 0.f;
  for (int t = 0; t < K; ++t) {
    const int qs = merge_pick(hv);
    const int32_t nn = hv[qs];
    float ab, de, gh;
    weights(x[nn], y[nn], z[nn], ab, de, gh);
    r = cic_add(r, corner_term(qs, c, ab, de, gh));
    merge_advance(qs, s, e, hv, idx);
  }
  rho[cid] = r;
}

// One wave per long cell: its contributions' terms formed in merged order
// into LDS, LC_CHUNK at a time, and added in order (wave_add_step).  When
// the eight lists hold disjoint index ranges -- the particles of a base
// cell are contiguous when the input is in cell order, as the resort leaves
// it -- the merge is the lists in the order of their first particles, and
// all lanes form the terms; otherwise lane 0 merges them one contribution
// at a time, its merge state carried across chunks.  A wave per workgroup
// keeps many cells' latency-bound sums in flight on each CU.
constexpr int LC_CHUNK = 2 * 64 * LT;

__global__ void __launch_bounds__(64)
k_cic_long_sum(Grid g, Ext ex, const int32_t *__restrict__ beg, const int32_t *__restrict__ end,
               const int32_t *__restrict__ idx, const uint32_t *__restrict__ long_cell,
               const int64_t *__restrict__ long_k, const unsigned long long *__restrict__ nlong,
               const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
               float c, float *__restrict__ rho) {
  __shared__ double tb[LC_CHUNK];
  __shared__ int32_t ss[8], se[8], soff[8];
  __shared__ int disjoint;
  const int tid = threadIdx.x;
  const int64_t nl = (int64_t)*nlong;
  for (int64_t t = blockIdx.x; t < nl; t += gridDim.x) {
    const uint32_t cid = long_cell[t];
    const int64_t K = long_k[t];
    int32_t s[8], e[8], hv[8];  // lane 0's merge state
    if (tid == 0) {
      const int k = (int)(cid % (uint32_t)g.ng2);
      const uint32_t ij = cid / (uint32_t)g.ng2;
      const int j = (int)(ij % (uint32_t)g.ng1), i = (int)(ij / (uint32_t)g.ng1);
      cell_lists(i, j, k, ex, beg, end, s, e);
      // each non-empty list's first and last particle; the lists ordered by
      // their first (a selection sort of eight)
      int32_t first[8], last[8], ord[8];
      int nq = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        ss[q] = s[q];
        se[q] = e[q];
        soff[q] = 0;
        first[q] = s[q] < e[q] ? idx[s[q]] : INT_MAX;
        last[q] = s[q] < e[q] ? idx[e[q] - 1] : INT_MAX;
        hv[q] = first[q];
      }
      for (int q = 0; q < 8; ++q)
        if (s[q] < e[q]) ord[nq++] = q;
      for (int a2 = 0; a2 < nq; ++a2)
        for (int b2 = a2 + 1; b2 < nq; ++b2)
          if (first[ord[b2]] < first[ord[a2]]) {
            const int tq = ord[a2];
            ord[a2] = ord[b2];
            ord[b2] = tq;
          }
      int dj = 1, off = 0;
      for (int a2 = 0; a2 < nq; ++a2) {
        const int q = ord[a2];
        if (a2 + 1 < nq && !(last[q] < first[ord[a2 + 1]])) dj = 0;
        soff[q] = off;
        off += e[q] - s[q];
      }
      disjoint = dj;
    }
    __syncthreads();
    float r = 0.f;
    for (int64_t c0 = 0; c0 < K; c0 += LC_CHUNK) {
      const int cl = (int)min<int64_t>(LC_CHUNK, K - c0);
      if (disjoint) {
        for (int u = tid; u < cl; u += blockDim.x) {
          const int pos = (int)c0 + u;
          // the list whose merged range holds pos (the ranges tile [0, K))
          int q = -1;
#pragma unroll
          for (int qq = 0; qq < 8; ++qq)
            if (pos >= soff[qq] && pos < soff[qq] + (se[qq] - ss[qq])) q = qq;
          double term = 0.0;
          if (q >= 0) {
            const int32_t nn = idx[ss[q] + (pos - soff[q])];
            float ab, de, gh;
            weights(x[nn], y[nn], z[nn], ab, de, gh);
            term = corner_term(q, c, ab, de, gh);
          }
          tb[u] = term;
        }
      } else if (tid == 0) {
        for (int u = 0; u < cl; ++u) {
          const int qs = merge_pick(hv);
          const int32_t nn = hv[qs];
          float ab, de, gh;
          weights(x[nn], y[nn], z[nn], ab, de, gh);
          tb[u] = corner_term(qs, c, ab, de, gh);
          merge_advance(qs, s, e, hv, idx);
        }
      }
      __syncthreads();
      {
        for (int b = 0; b < cl; b += 64 * LT) {
          const int cnt = min(64 * LT, cl - b);
          double tt[LT];
#pragma unroll
          for (int k = 0; k < LT; ++k) {
            const int i = LT * tid + k;
            tt[k] = i < cnt ? tb[b + i] : 0.0;
          }
          r = wave_add_step(r, tt, cnt);
        }
      }
      __syncthreads();  // tb is rewritten by the next chunk
    }
    if (tid == 0) rho[cid] = r;
  }
}

// The outside slot's contributions in (particle, corner) order, from
// position base (out_base: exclusive prefix of the particles' counts).
__global__ void __launch_bounds__(256)
k_cic_out_terms(int64_t n, const float *__restrict__ x, const float *__restrict__ y,
                const float *__restrict__ z, Grid g, float c, const int32_t *__restrict__ nout,
                const int32_t *__restrict__ out_base, int64_t base, uint32_t *__restrict__ skey,
                double *__restrict__ ts) {
  const int64_t nn = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (nn >= n || nout[nn] == 0) return;
  const Corners cc = corners(x[nn], y[nn], z[nn], g);
  int64_t p = base + out_base[nn];
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (cc.cell[q] == g.safe) {
      skey[p] = g.safe;
      ts[p] = corner_term(q, c, cc.ab, cc.de, cc.gh);
      ++p;
    }
}


// Particles.cxx:682-711, one component: f accumulated as the deposit's runs
// are (grid[safe] read as 0), then v += f * fscal * tau in float.
__global__ void __launch_bounds__(256)
k_inverse_cic(int64_t n, const float *__restrict__ x, const float *__restrict__ y,
              const float *__restrict__ z, Grid g, const float *__restrict__ grid,
              float tau, float fscal, float *__restrict__ v) {
  const int64_t nn = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (nn >= n) return;
  const Corners cc = corners(x[nn], y[nn], z[nn], g);
  float f = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const float s = cc.cell[q] == g.safe ? 0.f : grid[cc.cell[q]];
    const double t = corner_term(q, s, cc.ab, cc.de, cc.gh);
    if (q == 0) f = f + (float)t;
    else f = (float)((double)f + t);
  }
  v[nn] = v[nn] + f * fscal * tau;
}

__global__ void k_zero_u64(unsigned long long *p) {
  if (threadIdx.x == 0) p[0] = 0ull;
}

__global__ void k_fill0(int64_t n, float *a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0.f;
}

Grid make_grid(const int32_t ng[3]) {
  Grid g;
  g.ng0 = ng[0];
  g.ng1 = ng[1];
  g.ng2 = ng[2];
  g.safe = (uint32_t)((uint64_t)ng[0] * ng[1] * ng[2]);
  return g;
}

}  // namespace

namespace {

__global__ void k_fill_i32(int64_t n, int32_t *a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0;
}

template <class T>
hsr_status pm_excl_scan(hsr_ctx *c, const T *in, T *out, int64_t n) {
  size_t tmp = 0;
  HSR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, c->stream));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp + 256));
  HSR_HIP(hipcub::DeviceScan::ExclusiveSum(c->s.cub.p, tmp, in, out, (int)n, c->stream));
  return HSR_OK;
}

}  // namespace

hsr_status cic_run(hsr_ctx *c, int64_t n, const float *x, const float *y, const float *z,
                   const int32_t ng[3], float gpscal, float *rho) {
  hipStream_t st = c->stream;
  const Grid g = make_grid(ng);
  const int64_t ncell = (int64_t)g.safe + 1;
  const uint64_t ext = (uint64_t)(ng[0] + 1) * (uint64_t)(ng[1] + 1) * (uint64_t)(ng[2] + 1);
  if (ext >= 0xffffffffull) {
    hsr_set_error("hsr_cic: (ng[0]+1)(ng[1]+1)(ng[2]+1) must be below 2^32 - 1");
    return HSR_ERR_ARG;
  }
  k_fill0<<<nblk(ncell, 256), 256, 0, st>>>(ncell, rho);
  HSR_HIP(hipGetLastError());
  if (n <= 0) return HSR_OK;
  const float cc = gpscal * gpscal * gpscal;  // Particles.cxx:606
  Ext ex;
  ex.e1 = (uint32_t)ng[1] + 1u;
  ex.e2 = (uint32_t)ng[2] + 1u;
  ex.out = (uint32_t)ext;
  int bits = 1;
  while (bits < 32 && ((uint64_t)ex.out >> bits) != 0) ++bits;
  // layout (8-byte arrays first): per grid cell the long-cell count (8 B);
  // per staged contribution of the outside slot its term (8 B; at most 8
  // per particle); then 4-byte arrays: per particle its key, sorted key,
  // index, sorted index, outside count and its prefix; per base cell its
  // range; per grid cell the long cell's id; per staged contribution its
  // run key; the long-cell counter
  const size_t nn = (size_t)n, nc = (size_t)g.safe, ms = 8 * nn, ne = (size_t)ex.out;
  // (each array starts on a 256-byte boundary: 14 arrays)
  const size_t bytes = 8 * (nc + ms) + 4 * (6 * nn + 2 * ne + nc + ms) + 8 + 14 * 256;
  HSR_HIP(dbuf_reserve(c->sub.keys, bytes));
  char *bp = (char *)c->sub.keys.p;
  auto take = [&bp](size_t b) {
    char *r = bp;
    bp += (b + 255) & ~(size_t)255;
    return r;
  };
  int64_t *long_k = (int64_t *)take(8 * nc);
  double *ts = (double *)take(8 * ms);
  uint32_t *key = (uint32_t *)take(4 * nn);
  uint32_t *key2 = (uint32_t *)take(4 * nn);
  int32_t *idx = (int32_t *)take(4 * nn);
  int32_t *idx2 = (int32_t *)take(4 * nn);
  int32_t *nout = (int32_t *)take(4 * nn);
  int32_t *out_base = (int32_t *)take(4 * nn);
  int32_t *beg = (int32_t *)take(4 * ne);
  int32_t *end = (int32_t *)take(4 * ne);
  uint32_t *long_cell = (uint32_t *)take(4 * nc);
  uint32_t *skey = (uint32_t *)take(4 * ms);
  unsigned long long *nlong = (unsigned long long *)take(8);
  if ((size_t)(bp - (char *)c->sub.keys.p) > c->sub.keys.bytes) {
    hsr_set_error("hsr_cic: staging layout exceeds its buffer");
    return HSR_ERR_STATE;
  }
  k_zero_u64<<<1, 64, 0, st>>>(nlong);
  // particles by base cell (stable)
  k_cic_base<<<nblk(n, 256), 256, 0, st>>>(n, x, y, z, g, ex, key, idx, nout);
  size_t tmp = 0;
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key2, idx, idx2, (int)n, 0, bits, st));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp + 256));
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(c->s.cub.p, tmp, key, key2, idx, idx2, (int)n, 0, bits, st));
  k_fill_i32<<<nblk((int64_t)ne, 256), 256, 0, st>>>((int64_t)ne, beg);
  k_fill_i32<<<nblk((int64_t)ne, 256), 256, 0, st>>>((int64_t)ne, end);
  k_cic_ranges<<<nblk(n, 256), 256, 0, st>>>(n, key2, ex, beg, end);
  // short cells summed, long ones queued and summed a workgroup each (the
  // grid loops over the device-side count)
  k_cic_cells<<<nblk((int64_t)nc, 256), 256, 0, st>>>(g, ex, beg, end, idx2, x, y, z, cc, rho, long_cell,
                                                      long_k, nlong);
  const unsigned lgrid = (unsigned)std::max(1, c->num_cus) * 32;
  k_cic_long_sum<<<lgrid, 64, 0, st>>>(g, ex, beg, end, idx2, long_cell, long_k, nlong, x, y, z, cc, rho);
  // the outside slot: its contributions in (particle, corner) order
  hsr_status s;
  if ((s = pm_excl_scan(c, nout, out_base, n)) != HSR_OK) return s;
  int32_t last[2] = {0, 0};
  HSR_HIP(hipMemcpyAsync(&last[0], out_base + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipMemcpyAsync(&last[1], nout + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(stream_wait(st));
  const int64_t m = (int64_t)last[0] + last[1];
  if ((size_t)m > ms) {
    hsr_set_error("hsr_cic: staged contributions exceed 8 per particle");
    return HSR_ERR_STATE;
  }
  if (m > 0) {
    k_cic_out_terms<<<nblk(n, 256), 256, 0, st>>>(n, x, y, z, g, cc, nout, out_base, 0, skey, ts);
    k_cic_tail<<<1, TAIL_THREADS, 0, st>>>(m, skey, ts, rho);
  }
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

hsr_status inverse_cic_run(hsr_ctx *c, int64_t n, const float *x, const float *y,
                           const float *z, const int32_t ng[3], const float *grid, float tau,
                           float fscal, float *v) {
  if (n <= 0) return HSR_OK;
  const Grid g = make_grid(ng);
  k_inverse_cic<<<nblk(n, 256), 256, 0, c->stream>>>(n, x, y, z, g, grid, tau, fscal, v);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}
