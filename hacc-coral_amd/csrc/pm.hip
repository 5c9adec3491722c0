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
// in parallel it writes the 8 contributions of every particle (cell, term)
// in (particle, corner) order, sorts them stably by cell, and adds each
// cell's run in that order: one thread per short run, one wave per long run
// (a dense clump), streaming the run 64 contributions at a time through
// exact fixed-grid maps of the float accumulation.  The interpolation is a
// gather, one thread per particle.
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

// Contributions of particle nn at 8 nn + q: its corner cell (sort key), its
// position in the sequential order (sort value) and its term.
__global__ void __launch_bounds__(256)
k_cic_contrib(int64_t n, const float *__restrict__ x, const float *__restrict__ y,
              const float *__restrict__ z, Grid g, float c, uint32_t *__restrict__ key,
              int32_t *__restrict__ val, double *__restrict__ term) {
  const int64_t nn = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (nn >= n) return;
  const Corners cc = corners(x[nn], y[nn], z[nn], g);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int64_t k = 8 * nn + q;
    key[k] = cc.cell[q];
    val[k] = (int32_t)k;
    term[k] = corner_term(q, c, cc.ab, cc.de, cc.gh);
  }
}

// Terms in sorted order.
__global__ void __launch_bounds__(256)
k_cic_gather(int64_t m, const int32_t *__restrict__ val, const double *__restrict__ term,
             double *__restrict__ ts) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= m) return;
  ts[p] = term[val[p]];
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

constexpr int RUN_SHORT = 64;

// One thread per run of equal cells in the sorted contributions: runs of at
// most RUN_SHORT are added in order here; longer runs are queued for a wave
// each, except the last run of the sorted order, which is the outside slot
// whenever any corner fell outside the grid -- typically the longest run by
// far -- and goes to a workgroup of its own.  Each wave owns 64 positions
// and stages their keys and terms in LDS with coalesced loads; a head
// thread adds its run from LDS, and only the one run that crosses the end
// of the window reads its remaining terms from memory.
constexpr int RUNS_WAVES = 4;
__global__ void __launch_bounds__(64 * RUNS_WAVES)
k_cic_runs(int64_t m, const uint32_t *__restrict__ key, const double *__restrict__ ts,
           float *__restrict__ rho, int64_t *__restrict__ long_runs,
           unsigned long long *__restrict__ nlong, int64_t *__restrict__ tail_start) {
  __shared__ uint32_t sk[RUNS_WAVES][64];
  __shared__ double sv[RUNS_WAVES][64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)blockIdx.x * RUNS_WAVES + w) * 64;
  const int64_t p = w0 + l;
  // past the end: a key no cell has (cells and the outside slot are below
  // 2^31)
  const uint32_t cell = p < m ? key[p] : 0xffffffffu;
  sk[w][l] = cell;
  sv[w][l] = p < m ? ts[p] : 0.0;
  __syncthreads();
  // run ends in the window, by ballot: lane l's run ends at the first end at
  // or after l
  const bool endm = l == 63 || sk[w][l + 1 < 64 ? l + 1 : 63] != cell;
  const uint64_t ends = __ballot(endm);
  if (p >= m) return;
  if (p > 0 && (l > 0 ? sk[w][l - 1] : key[p - 1]) == cell) return;
  const int e = l + 1 + __builtin_ctzll(ends >> l);
  int64_t eg = w0 + e;  // the run continues past the window: from memory
  if (e == 64)
    while (eg < m && eg < p + RUN_SHORT + 1 && key[eg] == cell) ++eg;
  if (eg - p > RUN_SHORT) {
    if (key[m - 1] == cell) {
      // the last run (the outside slot, or the largest cell): a workgroup
      *tail_start = p;
      return;
    }
    const unsigned long long slot = atomicAdd(nlong, 1ull);
    long_runs[slot] = p;
    return;
  }
  // in order, eight terms read from LDS at a time
  float r = 0.f;
  for (int q = l; q < e; q += 8) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = sv[w][min(q + k, 63)];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (q + k < e) r = cic_add(r, v[k]);
  }
  for (int64_t q = w0 + 64; q < eg; ++q) r = cic_add(r, ts[q]);
  rho[cell] = r;
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

// One wave per long run: the lanes load LT consecutive contributions each,
// 64 LT at a time.  With all terms nonnegative (the deposit's weights are)
// the wave applies them by the map scan above -- each lane composes its LT
// terms in order, then the lanes' maps are scanned -- and applies the
// longest prefix that stays in the binade; in the lane where the sum leaves
// it the terms are stepped one at a time to the crossing term, which is
// added directly.  Otherwise the terms are added in order, reading them
// with v_readlane.  Every lane carries the same sum.  Waves loop over the
// queue; each exits when the queue is drained.
constexpr int LT = 4;

__device__ __forceinline__ double pick(const double (&t)[LT], int k) {
  return k == 0 ? t[0] : k == 1 ? t[1] : k == 2 ? t[2] : t[3];
}

__global__ void __launch_bounds__(64)
k_cic_long(int64_t m, const uint32_t *__restrict__ key, const double *__restrict__ ts,
           float *__restrict__ rho, const int64_t *__restrict__ long_runs,
           const unsigned long long *__restrict__ nlong) {
  static_assert(LT == 4, "pick() reads four terms per lane");
  const int l = threadIdx.x;
  const int64_t nruns = (int64_t)*nlong;
  for (int64_t run = blockIdx.x; run < nruns; run += gridDim.x) {
    const int64_t p0 = long_runs[run];
    const uint32_t cell = key[p0];
    float r = 0.f;
    bool done = false;
    for (int64_t base = p0; !done; base += 64 * LT) {
      double t[LT];
      int mine = 0;
      bool neg = false;
#pragma unroll
      for (int k = 0; k < LT; ++k) {
        const int64_t q = base + LT * l + k;
        const bool in = q < m && key[q] == cell;
        t[k] = in ? ts[q] : 0.0;
        mine += in;
        neg = neg || (in && !(t[k] >= 0.0));
      }
      // contributions of one cell are contiguous: the live terms are a prefix
      int cnt = mine;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
      done = cnt < 64 * LT;
      if (__ballot(neg) != 0ull) {
        for (int i = 0; i < cnt; ++i) r = cic_add(r, readlane_f64(pick(t, i % LT), i / LT));
        continue;
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
    }
    if (l == 0) rho[cell] = r;
  }
}

constexpr int TAIL_THREADS = 1024;

// The last run [p0, m) by one workgroup: LT TAIL_THREADS terms per step, LT
// per thread composed in order, maps scanned within each wave and across
// the 16 waves through LDS, the longest prefix that stays in the binade
// applied; the thread holding the crossing term steps its terms to it and
// adds it directly.
__global__ void __launch_bounds__(TAIL_THREADS)
k_cic_tail(int64_t m, const uint32_t *__restrict__ key, const double *__restrict__ ts,
           float *__restrict__ rho, const int64_t *__restrict__ tail_start) {
  __shared__ PMap wtot[TAIL_THREADS / 64];
  __shared__ float r_sh;
  __shared__ int j_sh, neg_sh, s_sh, rj_sh;
  const int64_t p0 = *tail_start;
  if (p0 < 0) return;
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

__global__ void k_zero_u64(unsigned long long *p, int64_t *tail) {
  if (threadIdx.x == 0) {
    p[0] = 0ull;
    tail[0] = -1;
  }
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

hsr_status cic_run(hsr_ctx *c, int64_t n, const float *x, const float *y, const float *z,
                   const int32_t ng[3], float gpscal, float *rho) {
  hipStream_t st = c->stream;
  const Grid g = make_grid(ng);
  const int64_t ncell = (int64_t)g.safe + 1;
  k_fill0<<<nblk(ncell, 256), 256, 0, st>>>(ncell, rho);
  HSR_HIP(hipGetLastError());
  if (n <= 0) return HSR_OK;
  const float cc = gpscal * gpscal * gpscal;  // Particles.cxx:606
  const int64_t m = 8 * n;
  int bits = 1;
  while (bits < 32 && ((uint64_t)g.safe >> bits) != 0) ++bits;
  // per contribution: key, key', val, val' (4 x 4 B), its term and the
  // sorted terms (2 x 8 B) and the long-run queue (8 B; at most one run per
  // contribution)
  const size_t mm = (size_t)m;
  HSR_HIP(dbuf_reserve(c->sub.keys, 40 * mm + 1024));
  uint32_t *key = dptr<uint32_t>(c->sub.keys);
  uint32_t *key2 = key + mm;
  int32_t *val = (int32_t *)(key2 + mm);
  int32_t *val2 = val + mm;
  double *term = (double *)(val2 + mm);  // 16 m bytes in: 8-byte aligned
  double *ts = term + mm;
  int64_t *long_runs = (int64_t *)(ts + mm);
  unsigned long long *nlong = (unsigned long long *)(long_runs + mm);
  int64_t *tail_start = (int64_t *)(nlong + 1);
  k_zero_u64<<<1, 64, 0, st>>>(nlong, tail_start);
  k_cic_contrib<<<nblk(n, 256), 256, 0, st>>>(n, x, y, z, g, cc, key, val, term);
  HSR_HIP(hipGetLastError());
  size_t tmp = 0;
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key2, val, val2, (int)m, 0, bits, st));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp));
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(c->s.cub.p, tmp, key, key2, val, val2, (int)m, 0, bits, st));
  k_cic_gather<<<nblk(m, 256), 256, 0, st>>>(m, val2, term, ts);
  k_cic_runs<<<nblk(m, 64 * RUNS_WAVES), 64 * RUNS_WAVES, 0, st>>>(m, key2, ts, rho, long_runs, nlong,
                                                                  tail_start);
  const int waves = std::max(1, c->num_cus) * 32;  // latency-bound: many waves in flight
  k_cic_long<<<waves, 64, 0, st>>>(m, key2, ts, rho, long_runs, nlong);
  k_cic_tail<<<1, TAIL_THREADS, 0, st>>>(m, key2, ts, rho, tail_start);
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
