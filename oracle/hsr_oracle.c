/*
 * hsr_oracle.c -- CPU restatement of HACC's short-range RCB force path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 * Build: oracle/Makefile, gcc -O2 -ffp-contract=off -fopenmp.  The
 * reference's x86 build has no -march flag (src/env/bashrc.linux-generic.gcc
 * :4-16), so it never contracts a*b+c into an FMA; -ffp-contract=off keeps
 * that operation sequence here.
 */
#include "hsr_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* OpenMP threads of the parallel loops below (the CPU baseline times 1 and
 * all threads); returns the previous setting. */
int hsro_set_threads(int n) {
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
}

/* ------------------------------------------------------------------ cm */
/* BGQCM.c:181-212 (generic branch): tight bbox via fminf/fmaxf seeded with
 * element 0; centre of mass as sequential double sums of float products,
 * divided in double and rounded to float. */
void hsro_cm(int64_t count, const float *xx, const float *yy, const float *zz,
             const float *mass, float *xmin, float *xmax, float *xc) {
  double x = 0, y = 0, z = 0, m = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (i == 0) {
      xmin[0] = xmax[0] = xx[0];
      xmin[1] = xmax[1] = yy[0];
      xmin[2] = xmax[2] = zz[0];
    } else {
      xmin[0] = fminf(xmin[0], xx[i]);
      xmax[0] = fmaxf(xmax[0], xx[i]);
      xmin[1] = fminf(xmin[1], yy[i]);
      xmax[1] = fmaxf(xmax[1], yy[i]);
      xmin[2] = fminf(xmin[2], zz[i]);
      xmax[2] = fmaxf(xmax[2], zz[i]);
    }
    float w = mass[i];
    x += (double)(w * xx[i]);
    y += (double)(w * yy[i]);
    z += (double)(w * zz[i]);
    m += (double)w;
  }
  xc[0] = (float)(x / m);
  xc[1] = (float)(y / m);
  xc[2] = (float)(z / m);
}

/* ----------------------------------------------------------- partition */
/* RCBForceTree.cxx:623-672.  idx collects ascending positions with
 * key < pv; then for j < is the elements at j and idx[j] are swapped in
 * every array.  Heap scratch instead of the reference's stack VLA. */
int64_t hsro_partition(int64_t n, float *xx, float *yy, float *zz, float *vx,
                       float *vy, float *vz, float *mass, float *phi,
                       int64_t *id, uint16_t *mask, float pv) {
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t is = 0;
  for (int64_t i = 0; i < n; ++i)
    if (xx[i] < pv) idx[is++] = i;
  for (int64_t j = 0; j < is; ++j) {
    int64_t i = idx[j];
    float t;
    int64_t t64;
    uint16_t t16;
    t = mass[i]; mass[i] = mass[j]; mass[j] = t;
    t = phi[i]; phi[i] = phi[j]; phi[j] = t;
    t16 = mask[i]; mask[i] = mask[j]; mask[j] = t16;
    t64 = id[i]; id[i] = id[j]; id[j] = t64;
    t = xx[i]; xx[i] = xx[j]; xx[j] = t;
    t = yy[i]; yy[i] = yy[j]; yy[j] = t;
    t = zz[i]; zz[i] = zz[j]; zz[j] = t;
    t = vx[i]; vx[i] = vx[j]; vx[j] = t;
    t = vy[i]; vy[i] = vy[j]; vy[j] = t;
    t = vz[i]; vz[i] = vz[j]; vz[j] = t;
  }
  free(idx);
  return is;
}

/* ------------------------------------------------------- pseudo-particles */
/* RCBForceTree.cxx:519-523 */
static float pptdr(const float *xmin, const float *xmax, const float *xc) {
  float a = xmax[0] - xc[0], b = xmax[1] - xc[1], c = xmax[2] - xc[2];
  float d = xc[0] - xmin[0], e = xc[1] - xmin[1], f = xc[2] - xmin[2];
  /* std::min(a, std::min(b, std::min(c, std::min(d, std::min(e, f))))) ;
   * std::min(p, q) returns (q < p) ? q : p. */
  float r = (f < e) ? f : e;
  r = (r < d) ? r : d;
  r = (r < c) ? r : c;
  r = (r < b) ? r : b;
  r = (r < a) ? r : a;
  return r;
}

/* The spherical designs of RCBForceTree.cxx:134-272 that this path uses:
 * TDPTS = 1, the single point (0,0,0) (the monopole at the centre of mass);
 * TDPTS = 12, Hardin and Sloane's 12-point 4-design (icosahedron vertices),
 * in the reference's order, rounded to float as POSVEL_T arrays are. */
static const float SPH12_X[12] = {
    0.0f, 0.0f, (float)0.525731112119134, (float)-0.525731112119134,
    (float)0.85065080835204, (float)-0.85065080835204, 0.0f, 0.0f,
    (float)-0.525731112119134, (float)0.525731112119134,
    (float)-0.85065080835204, (float)0.85065080835204};
static const float SPH12_Y[12] = {
    (float)0.85065080835204, (float)0.85065080835204, 0.0f, 0.0f,
    (float)0.525731112119134, (float)0.525731112119134,
    (float)-0.85065080835204, (float)-0.85065080835204, 0.0f, 0.0f,
    (float)-0.525731112119134, (float)-0.525731112119134};
static const float SPH12_Z[12] = {
    (float)0.525731112119134, (float)-0.525731112119134,
    (float)0.85065080835204, (float)0.85065080835204, 0.0f, 0.0f,
    (float)-0.525731112119134, (float)0.525731112119134,
    (float)-0.85065080835204, (float)-0.85065080835204, 0.0f, 0.0f};

static int tdpts_of(const hsro_params *p) { return p->tdpts > 1 ? p->tdpts : 1; }

/* The pseudo masses of node k. */
static float *ppm_of(const hsro_params *p, hsro_node *t, int64_t k) {
  return tdpts_of(p) == 1 ? t[k].ppm : p->ppm_t + k * (int64_t)tdpts_of(p);
}
static const float *ppm_of_c(const hsro_params *p, const hsro_node *t, int64_t k) {
  return tdpts_of(p) == 1 ? t[k].ppm : p->ppm_t + k * (int64_t)tdpts_of(p);
}

/* RCBForceTree.cxx:525-534: ppx[i] = tdr*sphdesign<T>::x[i] + xc[0]. */
static void pppts(int T, float tdr, const float *xc, float *px, float *py, float *pz) {
  for (int i = 0; i < T; ++i) {
    float sx = T == 1 ? 0.0f : SPH12_X[i];
    float sy = T == 1 ? 0.0f : SPH12_Y[i];
    float sz = T == 1 ? 0.0f : SPH12_Z[i];
    px[i] = tdr * sx + xc[0];
    py[i] = tdr * sy + xc[1];
    pz[i] = tdr * sz + xc[2];
  }
}

void hsro_pppts(int tdpts, float tdr, const float *xc, float *px, float *py,
                float *pz) {
  pppts(tdpts > 1 ? tdpts : 1, tdr, xc, px, py, pz);
}

/* RCBForceTree.cxx:536-569: ppm[j] += mass[i]*(odr0 + odr1 + odr2) with
 * K = T, odr0 = 1/K, odr1 = (3/K)(ri/tdr) aij and
 * odr2 = (5/K)(ri/tdr)(ri/tdr) 0.5 (3 aij^2 - 1), the product with 0.5
 * taken in double as the double literal makes it; odr1 = odr2 = 0 where the
 * design point sits at the centre (T = 1). */
static void ppT(int T, int64_t count, const float *xx, const float *yy,
                const float *zz, const float *mass, const float *xc,
                const float *ppx, const float *ppy, const float *ppz,
                float *ppm, float tdr) {
  const float K = (float)T;
  const float odr0 = 1 / K;
  for (int64_t i = 0; i < count; ++i) {
    float xi = xx[i] - xc[0], yi = yy[i] - xc[1], zi = zz[i] - xc[2];
    float ri = sqrtf(xi * xi + yi * yi + zi * zi);
    for (int j = 0; j < T; ++j) {
      float xj = ppx[j] - xc[0], yj = ppy[j] - xc[1], zj = ppz[j] - xc[2];
      float rj2 = xj * xj + yj * yj + zj * zj;
      float odr1 = 0, odr2 = 0;
      if (rj2 != 0) {
        float rj = sqrtf(rj2);
        float aij = (xi * xj + yi * yj + zi * zj) / (ri * rj);
        odr1 = (3 / K) * (ri / tdr) * aij;
        odr2 = (float)((double)((5 / K) * (ri / tdr) * (ri / tdr)) * 0.5 *
                       (double)(3 * aij * aij - 1));
      }
      ppm[j] += mass[i] * (odr0 + odr1 + odr2);
    }
  }
}

/* ---------------------------------------------------------------- build */
typedef struct {
  const hsro_params *p;
  float *a[8]; /* x y z vx vy vz mass phi */
  int64_t *id;
  uint16_t *mask;
  hsro_node *t;
  int64_t size, cap;
  int overflow;
} builder;

static void build_node(builder *b, int64_t tl);

/* RCBForceTree.cxx:690-775 (unthreaded): rotate coordinate pointers by d,
 * partition at xc[d], and recurse left then right. */
static void build_subtree(builder *b, int d, int64_t tl, int64_t tlcl,
                          int64_t tlcr) {
  float *x1, *x2, *x3;
  float *xx = b->a[0], *yy = b->a[1], *zz = b->a[2];
  if (d == 0) { x1 = xx; x2 = yy; x3 = zz; }
  else if (d == 1) { x1 = yy; x2 = zz; x3 = xx; }
  else { x1 = zz; x2 = xx; x3 = yy; }
  hsro_node *t = b->t;
  float split = t[tl].xc[d];
  int64_t off = t[tl].offset;
  int64_t is = hsro_partition(t[tl].count, x1 + off, x2 + off, x3 + off,
                              b->a[3] + off, b->a[4] + off, b->a[5] + off,
                              b->a[6] + off, b->a[7] + off, b->id + off,
                              b->mask + off, split);
  if (is == 0 || is == t[tl].count) return;
  t[tlcl].count = is;
  t[tlcr].count = t[tl].count - t[tlcl].count;
  if (t[tlcl].count > 0) {
    t[tl].cl = tlcl;
    t[tlcl].offset = t[tl].offset;
    t[tlcl].xmax[d] = split;
    build_node(b, tlcl);
  }
  t = b->t;
  if (t[tlcr].count > 0) {
    t[tl].cr = tlcr;
    t[tlcr].offset = t[tl].offset + t[tlcl].count;
    t[tlcr].xmin[d] = split;
    build_node(b, tlcr);
  }
}

/* RCBForceTree.cxx:779-890 */
static void build_node(builder *b, int64_t tl) {
  hsro_node *t = b->t;
  int64_t cnt = t[tl].count, off = t[tl].offset;
  float **a = b->a;
  const int T = tdpts_of(b->p);
  hsro_cm(cnt, a[0] + off, a[1] + off, a[2] + off, a[6] + off, t[tl].xmin,
          t[tl].xmax, t[tl].xc);
  if (cnt <= b->p->nd) {
    t[tl].tdr = b->p->ppc * pptdr(t[tl].xmin, t[tl].xmax, t[tl].xc);
    float *pm = ppm_of(b->p, t, tl);
    memset(pm, 0, sizeof(float) * T);
    if (cnt > T) {
      float px[12], py[12], pz[12];
      pppts(T, t[tl].tdr, t[tl].xc, px, py, pz);
      ppT(T, cnt, a[0] + off, a[1] + off, a[2] + off, a[6] + off, t[tl].xc,
          px, py, pz, pm, t[tl].tdr);
    }
    t[tl].ppm[0] = pm[0];
    return;
  }
  int64_t tlcl = b->size, tlcr = tlcl + 1;
  if (tlcr + 1 > b->cap) {
    b->overflow = 1;
    return;
  }
  b->size = tlcr + 1;
  memset(&t[tlcl], 0, sizeof(hsro_node) * 2);
  if (T > 1) memset(b->p->ppm_t + tlcl * (int64_t)T, 0, sizeof(float) * 2 * T);
  for (int i = 0; i < 3; ++i) {
    t[tlcl].xmin[i] = t[tl].xmin[i];
    t[tlcr].xmin[i] = t[tl].xmin[i];
    t[tlcl].xmax[i] = t[tl].xmax[i];
    t[tlcr].xmax[i] = t[tl].xmax[i];
  }
  float xlen[3];
  for (int i = 0; i < 3; ++i) xlen[i] = t[tl].xmax[i] - t[tl].xmin[i];
  int d;
  if (xlen[0] > xlen[1] && xlen[0] > xlen[2]) d = 0;
  else if (xlen[1] > xlen[2]) d = 1;
  else d = 2;
  build_subtree(b, d, tl, tlcl, tlcr);
  if (b->overflow) return;

  float ppx[12], ppy[12], ppz[12];
  t[tl].tdr = b->p->ppc * pptdr(t[tl].xmin, t[tl].xmax, t[tl].xc);
  pppts(T, t[tl].tdr, t[tl].xc, ppx, ppy, ppz);
  float *pm = ppm_of(b->p, t, tl);
  memset(pm, 0, sizeof(float) * T);
  const int64_t ch[2] = {tlcl, tlcr};
  for (int k = 0; k < 2; ++k) {
    int64_t c = ch[k];
    if (t[c].count > 0) {
      if (t[c].count <= T) {
        int64_t offc = t[c].offset;
        ppT(T, t[c].count, a[0] + offc, a[1] + offc, a[2] + offc, a[6] + offc,
            t[tl].xc, ppx, ppy, ppz, pm, t[tl].tdr);
      } else {
        float cx[12], cy[12], cz[12];
        pppts(T, t[c].tdr, t[c].xc, cx, cy, cz);
        ppT(T, T, cx, cy, cz, ppm_of(b->p, t, c), t[tl].xc, ppx, ppy, ppz, pm,
            t[tl].tdr);
      }
    }
  }
  t[tl].ppm[0] = pm[0];
}

int64_t hsro_build(const hsro_params *p, int64_t n, float *x, float *y,
                   float *z, float *vx, float *vy, float *vz, float *mass,
                   float *phi, int64_t *id, uint16_t *mask, hsro_node *nodes,
                   int64_t cap) {
  if (cap < 1) return -1;
  builder b;
  b.p = p;
  b.a[0] = x; b.a[1] = y; b.a[2] = z; b.a[3] = vx; b.a[4] = vy;
  b.a[5] = vz; b.a[6] = mass; b.a[7] = phi;
  b.id = id; b.mask = mask; b.t = nodes; b.cap = cap; b.size = 1;
  b.overflow = 0;
  if (!(p->tdpts == 0 || p->tdpts == 1 || p->tdpts == 12) ||
      (p->tdpts > 1 && !p->ppm_t))
    return -2;
  memset(&nodes[0], 0, sizeof(hsro_node));
  if (p->tdpts > 1) memset(p->ppm_t, 0, sizeof(float) * p->tdpts);
  nodes[0].count = n;
  nodes[0].offset = 0;
  for (int i = 0; i < 3; ++i) {
    nodes[0].xmin[i] = p->min_loc[i];
    nodes[0].xmax[i] = p->max_loc[i];
  }
  build_node(&b, 0);
  return b.overflow ? -1 : b.size;
}

/* ------------------------------------------------------------ leaf order */
/* RCBForceTree.cxx:1151-1197.  out receives, for every leaf inside the
 * force box, its node id; the leaves come in the task-spawn order. */
int64_t hsro_leaf_order(const hsro_params *p, const hsro_node *t,
                        int64_t nnodes, int64_t *out, int64_t cap) {
  int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (nnodes + 2));
  int64_t *parents = (int64_t *)malloc(sizeof(int64_t) * (nnodes + 2));
  int64_t nq = 0, np = 0, nout = 0;
  q[nq++] = 0;
  while (nq > 0) {
    int64_t tl = q[nq - 1];
    if (t[tl].cr == 0 && t[tl].cl == 0) {
      --nq;
      int inside = 1;
      for (int i = 0; i < 3; ++i) {
        inside &= (t[tl].xmax[i] < p->max_force[i] &&
                   t[tl].xmax[i] > p->min_force[i]) ||
                  (t[tl].xmin[i] < p->max_force[i] &&
                   t[tl].xmin[i] > p->min_force[i]);
      }
      if (inside) {
        if (nout >= cap) { nout = -1; break; }
        out[nout++] = tl;
      }
    } else if (np > 0 && parents[np - 1] == tl) {
      --np;
      --nq;
    } else {
      if (t[tl].cl > 0) q[nq++] = t[tl].cl;
      if (t[tl].cr > 0) q[nq++] = t[tl].cr;
      parents[np++] = tl;
    }
  }
  free(q);
  free(parents);
  return nout;
}

/* ------------------------------------------------------------------ lists */
/* Ancestors of tl: walk from the root along the child whose particle range
 * contains tl's range (equivalent to the DFS `parents` stack at the time
 * tl is visited in calcInternodeForces). */
static int64_t ancestors(const hsro_node *t, int64_t tl, int64_t *par) {
  int64_t n = 0, cur = 0;
  int64_t off = t[tl].offset;
  while (cur != tl) {
    par[n++] = cur;
    int64_t cl = t[cur].cl, cr = t[cur].cr;
    if (cl == tl || cr == tl) break;
    if (cl > 0 && off >= t[cl].offset && off < t[cl].offset + t[cl].count)
      cur = cl;
    else
      cur = cr;
  }
  return n;
}

static int in_sorted(const int64_t *a, int64_t n, int64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo < n && a[lo] == v;
}

#define EMIT(A, B, L)                                \
  do {                                               \
    if (nseg >= cap) { ovf = 1; goto done; }         \
    seg_a[nseg] = (A);                               \
    seg_b[nseg] = (B);                               \
    ++nseg;                                          \
    nent += (L);                                     \
  } while (0)

/* RCBForceTree.cxx:924-1140 (useRealOA = false). */
int64_t hsro_list(const hsro_params *p, const hsro_node *t, int64_t nnodes,
                  int64_t tl, int64_t *seg_a, int64_t *seg_b, int64_t cap,
                  int64_t *n_entries) {
  const float fsrrmax = p->fsrrmax;
  const float fsrrmax2 = fsrrmax * fsrrmax;
  const float tanOA = p->tan_oa;
  int64_t *par = (int64_t *)malloc(sizeof(int64_t) * (nnodes + 1));
  int64_t npar = ancestors(t, tl, par);
  int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (nnodes + 2));
  int64_t nq = 0, nseg = 0, nent = 0;
  int ovf = 0;
  q[nq++] = 0;
  while (nq > 0) {
    int64_t tln = q[--nq];
    if (tln < tl) {
      if (in_sorted(par, npar, tln)) {
        int64_t tlncr = t[tln].cr, tlncl = t[tln].cl;
        if (tlncl != tl && tlncl > 0 && t[tlncl].count > 0) q[nq++] = tlncl;
        if (tlncr != tl && tlncr > 0 && t[tlncr].count > 0) q[nq++] = tlncr;
        continue;
      }
    }
    float dx = t[tln].xc[0] - t[tl].xc[0];
    float dy = t[tln].xc[1] - t[tl].xc[1];
    float dz = t[tln].xc[2] - t[tl].xc[2];
    float dist2 = dx * dx + dy * dy + dz * dz;
    float sx = t[tln].xmax[0] - t[tln].xmin[0];
    float sy = t[tln].xmax[1] - t[tln].xmin[1];
    float sz = t[tln].xmax[2] - t[tln].xmin[2];
    float syz = (sz * sz < sy * sy) ? sz * sz : sy * sy;
    float l2 = (syz < sx * sx) ? syz : sx * sx;
    float dtt2 = dist2 * tanOA * tanOA;
    int looksBig;
    if (l2 > dtt2) {
      looksBig = 1;
    } else {
      looksBig = 0;
      for (int i = 0; i < 2 && !looksBig; ++i)
        for (int j = 0; j < 2; ++j) {
          float x1 = (i == 0 ? t[tln].xmin : t[tln].xmax)[0] - t[tl].xc[0];
          float y1 = (j == 0 ? t[tln].xmin : t[tln].xmax)[1] - t[tl].xc[1];
          float z1 = t[tln].xmin[2] - t[tl].xc[2];
          float x2 = (i == 0 ? t[tln].xmax : t[tln].xmin)[0] - t[tl].xc[0];
          float y2 = (j == 0 ? t[tln].xmax : t[tln].xmin)[1] - t[tl].xc[1];
          float z2 = t[tln].xmax[2] - t[tl].xc[2];
          float ddx = x1 - x2, ddy = y1 - y2, ddz = z1 - z2;
          float dh2 = ddx * ddx + ddy * ddy + ddz * ddz;
          if (dh2 > dtt2) {
            looksBig = 1;
            break;
          }
        }
    }
    if (!looksBig) {
      if (dist2 > fsrrmax2) continue;
      if (t[tln].count <= tdpts_of(p)) {
        EMIT(t[tln].offset, t[tln].count, t[tln].count);
        continue;
      }
      EMIT(tln, -1, tdpts_of(p));
      continue;
    } else if (t[tln].cr == 0 && t[tln].cl == 0) {
      EMIT(t[tln].offset, t[tln].count, t[tln].count);
      continue;
    }
    int64_t ch[2] = {t[tln].cl, t[tln].cr};
    for (int k = 0; k < 2; ++k) {
      int64_t c = ch[k];
      if (c > 0 && t[c].count > 0) {
        int close = 1;
        for (int i = 0; i < 3; ++i) {
          float dist = 0;
          if (t[tl].xmax[i] < t[c].xmin[i])
            dist = t[c].xmin[i] - t[tl].xmax[i];
          else if (t[tl].xmin[i] > t[c].xmax[i])
            dist = t[tl].xmin[i] - t[c].xmax[i];
          if (dist > fsrrmax) {
            close = 0;
            break;
          }
        }
        if (close) q[nq++] = c;
      }
    }
  }
  EMIT(t[tl].offset, t[tl].count, t[tl].count);
done:
  free(par);
  free(q);
  if (n_entries) *n_entries = nent;
  return ovf ? -1 : nseg;
}

int64_t hsro_lists_all(const hsro_params *p, const hsro_node *t,
                       int64_t nnodes, const int64_t *leaves, int64_t nleaves,
                       int64_t *seg_ptr, int64_t *seg_a, int64_t *seg_b,
                       int64_t cap, int64_t *entries_per_leaf) {
  /* pass 1: counts (parallel), pass 2: fill */
  int64_t *nsegs = (int64_t *)malloc(sizeof(int64_t) * (nleaves + 1));
  int bad = 0;
#pragma omp parallel
  {
    int64_t lcap = 1 << 16;
    int64_t *la = (int64_t *)malloc(sizeof(int64_t) * lcap);
    int64_t *lb = (int64_t *)malloc(sizeof(int64_t) * lcap);
#pragma omp for schedule(dynamic, 16)
    for (int64_t k = 0; k < nleaves; ++k) {
      int64_t ne = 0;
      int64_t ns = hsro_list(p, t, nnodes, leaves[k], la, lb, lcap, &ne);
      while (ns < 0) {
        lcap *= 2;
        la = (int64_t *)realloc(la, sizeof(int64_t) * lcap);
        lb = (int64_t *)realloc(lb, sizeof(int64_t) * lcap);
        ns = hsro_list(p, t, nnodes, leaves[k], la, lb, lcap, &ne);
      }
      nsegs[k] = ns;
      if (entries_per_leaf) entries_per_leaf[k] = ne;
    }
    free(la);
    free(lb);
  }
  seg_ptr[0] = 0;
  for (int64_t k = 0; k < nleaves; ++k) seg_ptr[k + 1] = seg_ptr[k] + nsegs[k];
  free(nsegs);
  if (seg_ptr[nleaves] > cap) return -1;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < nleaves; ++k) {
    int64_t ne;
    int64_t ns = hsro_list(p, t, nnodes, leaves[k], seg_a + seg_ptr[k],
                           seg_b + seg_ptr[k], seg_ptr[k + 1] - seg_ptr[k],
                           &ne);
    if (ns != seg_ptr[k + 1] - seg_ptr[k]) bad = 1;
  }
  return bad ? -1 : seg_ptr[nleaves];
}

/* ------------------------------------------------------------ force laws */
/* BGQStep16.c:40-45,167 (also ForceLaw.cxx:109-114, POLY_ORDER 5). */
static const float P5[6] = {0.269327f, -0.0750978f, 0.0114808f,
                            -0.00109313f, 0.0000605491f, -0.00000147177f};
/* ForceLaw.cxx:98-104, POLY_ORDER 6 (stock). */
static const float P6[7] = {0.271431f, -0.0783394f, 0.0133122f,
                            -0.00159485f, 0.000132336f, -0.00000663394f,
                            0.000000147305f};
/* FGrid::m_rmax (ForceLaw.cxx:32) */
static const float FGRID_RMAX = 3.116326355f;

/* FGridEvalPoly::eval (ForceLaw.cxx:137-141) */
static float poly_eval(const float *a, int order, float r2) {
  float ret;
  if (order == 6)
    ret = a[0] + r2 * (a[1] + r2 * (a[2] + r2 * (a[3] + r2 * (a[4] +
          r2 * (a[5] + r2 * a[6])))));
  else
    ret = a[0] + r2 * (a[1] + r2 * (a[2] + r2 * (a[3] + r2 * (a[4] +
          r2 * (a[5] + r2 * 0.0f)))));
  float r2max = FGRID_RMAX * FGRID_RMAX;
  return ret * (float)(r2 >= 0.0f) * (float)(r2 <= r2max);
}

/* ForceLawSR::f_over_r (ForceLaw.cxx:187-192) */
static float sr_f_over_r(int order, float r2, float rsm2) {
  const float *a = order == 6 ? P6 : P5;
  float ret = powf(r2 + rsm2, -1.5f) - poly_eval(a, order, r2);
  float r2max = FGRID_RMAX * FGRID_RMAX;
  ret *= (float)(r2 >= 0.0f) * (float)(r2 <= r2max);
  return ret;
}

/* BGQStep16.c:172-182 scalar tail: f = pow(r2+rsm2, -1.5) - P5(r2) with the
 * power in double and the polynomial in float; masked by r2 < fsrrmax2 and
 * guarded by r2 > 0 (mass factor m_j applied by the caller). */
static float step16_f(float r2, float rsm2, float fsrrmax2) {
  float f = (float)(pow((double)(r2 + rsm2), -1.5) -
                    (double)(P5[0] + r2 * (P5[1] + r2 * (P5[2] + r2 * (P5[3] +
                             r2 * (P5[4] + r2 * P5[5]))))));
  float msk = (r2 < fsrrmax2) ? 1.0f : 0.0f;
  return (r2 > 0.0f) ? msk * f : 0.0f;
}

/* FGrid constants (ForceLaw.cxx:23-33). */
static const float FG_B = 0.72f, FG_C = 0.01f, FG_D = 0.27f, FG_E = 0.0001f,
                   FG_F = 360.0f, FG_G = 100.0f, FG_H = 0.67f, FG_L = 17.0f;

/* FGrid::fgor (ForceLaw.cxx:39-51): the fitted grid force over r, with the
 * reference's mix of float and double operations (C's promotion rules are
 * C++'s here). */
static float fgor(float r) {
  float f0 = FG_C + 2.0 / 3.0 * FG_B * FG_B * FG_B;
  float r2 = r * r;
  float r4 = r2 * r2;
  float r6 = r4 * r2;
  float coshbr = coshf(FG_B * r);
  float r3fgor = tanhf(FG_B * r) - FG_B * r / coshbr / coshbr +
                 FG_C * r * r2 * (1.0 + FG_D * r2) * expf(-1.0 * FG_D * r2) +
                 FG_E * r2 * (FG_F * r2 + FG_G * r4 + FG_L * r6) * expf(-1.0 * FG_H * r2);
  float rdiv = r + 1.0 * (r <= 0.0);
  float rdiv3 = rdiv * rdiv * rdiv;
  return (r3fgor / rdiv3 + (r <= 0.0) * f0) * (r <= FGRID_RMAX) * (r >= 0.0);
}

void hsro_fgrid_table(int n, float *r2, float *f) {
  double dr2 = (FGRID_RMAX * FGRID_RMAX) / (n - 1.0);
  for (int i = 0; i < n; i++) {
    r2[i] = i * dr2;
    f[i] = fgor(sqrt(i * dr2));
  }
}

/* FGridEvalInterp (ForceLaw.cxx:145-171) */
typedef struct {
  int n;
  float *r2, *f;
  float r2min, r2max, dr2, oodr2;
} interp_t;

static void interp_init(interp_t *I, int n) {
  I->n = n;
  I->r2 = (float *)malloc(sizeof(float) * n);
  I->f = (float *)malloc(sizeof(float) * n);
  hsro_fgrid_table(n, I->r2, I->f);
  I->r2min = I->r2[0];
  I->r2max = I->r2[n - 1];
  I->dr2 = (I->r2max - I->r2min) / (n - 1.0);
  I->oodr2 = 1.0 / I->dr2;
}

static void interp_free(interp_t *I) {
  free(I->r2);
  free(I->f);
}

static float interp_eval(const interp_t *I, float r2) {
  int inRange = (r2 > I->r2min) * (r2 < I->r2max);
  float inRangef = 1.0 * inRange;
  int indx = (int)((r2 - I->r2min) * I->oodr2) * inRange;
  /* the reference reads m_f[indx + 1] even when r2 rounds onto the last
   * node (indx = n - 1), one past its table; the oracle reads the last
   * entry there */
  int hi = indx + 1 < I->n ? indx + 1 : I->n - 1;
  return inRangef * (I->f[indx] + (r2 - I->r2[indx]) * I->oodr2 * (I->f[hi] - I->f[indx]));
}

/* ForceLawSR::f_over_r (ForceLaw.cxx:187-192) with FGridEvalFit
 * (r2max = rmax^2, ForceLaw.cxx:71-81) or FGridEvalInterp (r2min, r2max from
 * the table). */
static float sr_fit_f_over_r(float r2, float rsm2) {
  float r2max = FGRID_RMAX * FGRID_RMAX;
  float ret = powf(r2 + rsm2, -1.5f) - fgor(sqrtf(r2));
  ret *= (float)(r2 >= 0.0f) * (float)(r2 <= r2max);
  return ret;
}

static float sr_interp_f_over_r(const interp_t *I, float r2, float rsm2) {
  float ret = powf(r2 + rsm2, -1.5f) - interp_eval(I, r2);
  ret *= (float)(r2 >= I->r2min) * (float)(r2 <= I->r2max);
  return ret;
}

static float law_f32(int law, float r2, float rsm, float fsrrmax, const interp_t *I) {
  float rsm2 = rsm * rsm;
  switch (law) {
  case HSRO_LAW_STEP16: return step16_f(r2, rsm2, fsrrmax * fsrrmax);
  case HSRO_LAW_SR_POLY5: return sr_f_over_r(5, r2, rsm2);
  case HSRO_LAW_SR_POLY6: return sr_f_over_r(6, r2, rsm2);
  case HSRO_LAW_NEWTON: return (float)(1.0 / (double)r2 / sqrt((double)r2));
  case HSRO_LAW_SR_FIT: return sr_fit_f_over_r(r2, rsm2);
  case HSRO_LAW_SR_INTERP: return I ? sr_interp_f_over_r(I, r2, rsm2) : 0.0f;
  }
  return 0.0f;
}

float hsro_law_f32(int law, float r2, float rsm, float fsrrmax) {
  return law_f32(law, r2, rsm, fsrrmax, NULL);
}

void hsro_law_f32_vec(int law, int64_t n, const float *r2, float *out, float rsm,
                      float fsrrmax, int n_interp) {
  interp_t I;
  const int interp = law == HSRO_LAW_SR_INTERP && n_interp >= 2;
  if (interp) interp_init(&I, n_interp);
  for (int64_t k = 0; k < n; ++k)
    out[k] = law_f32(law, r2[k], rsm, fsrrmax, interp ? &I : NULL);
  if (interp) interp_free(&I);
}

int hsro_law_coeffs(int law, float *out) {
  switch (law) {
  case HSRO_LAW_STEP16:
  case HSRO_LAW_SR_POLY5:
    for (int k = 0; k < 6; ++k) out[k] = P5[k];
    return 6;
  case HSRO_LAW_SR_POLY6:
    for (int k = 0; k < 7; ++k) out[k] = P6[k];
    return 7;
  }
  return 0;
}

/* ------------------------------------------------------------- kernels */
/* Gather the list of one leaf into SoA (the reference's nx/ny/nz/nm). */
static int64_t gather_list(const hsro_params *p, const hsro_node *t,
                           const float *x, const float *y, const float *z,
                           const float *m, const int64_t *sa, const int64_t *sb,
                           int64_t ns, float **bufs, int64_t *bcap) {
  const int T = tdpts_of(p);
  int64_t len = 0;
  for (int64_t s = 0; s < ns; ++s) len += sb[s] >= 0 ? sb[s] : T;
  if (len > *bcap) {
    *bcap = len * 2;
    for (int k = 0; k < 4; ++k)
      bufs[k] = (float *)realloc(bufs[k], sizeof(float) * (*bcap));
  }
  int64_t o = 0;
  for (int64_t s = 0; s < ns; ++s) {
    if (sb[s] >= 0) {
      for (int64_t i = 0; i < sb[s]; ++i) {
        bufs[0][o] = x[sa[s] + i];
        bufs[1][o] = y[sa[s] + i];
        bufs[2][o] = z[sa[s] + i];
        bufs[3][o] = m[sa[s] + i];
        ++o;
      }
    } else {
      /* RCBForceTree.cxx:1051-1060: the node's T pseudo-particles */
      const hsro_node *n = &t[sa[s]];
      const float *pm = ppm_of_c(p, t, sa[s]);
      pppts(T, n->tdr, n->xc, &bufs[0][o], &bufs[1][o], &bufs[2][o]);
      for (int j = 0; j < T; ++j) bufs[3][o + j] = pm[j];
      o += T;
    }
  }
  return len;
}

void hsro_force_f32(int law, const hsro_params *p, const hsro_node *t,
                    const float *x, const float *y, const float *z,
                    const float *m, const int64_t *leaves, int64_t nleaves,
                    const int64_t *seg_ptr, const int64_t *seg_a,
                    const int64_t *seg_b, float *ax, float *ay, float *az) {
  const float fsrrmax2 = p->fsrrmax * p->fsrrmax;
  const float rsm2 = p->rsm * p->rsm;
  const float fcoeff = 1.0f;
  interp_t I;
  const int interp = law == HSRO_LAW_SR_INTERP && p->n_interp >= 2;
  if (interp) interp_init(&I, p->n_interp);
#pragma omp parallel
  {
    float *bufs[4] = {0, 0, 0, 0};
    int64_t bcap = 0;
#pragma omp for schedule(dynamic, 4)
    for (int64_t k = 0; k < nleaves; ++k) {
      const hsro_node *n = &t[leaves[k]];
      int64_t len = gather_list(p, t, x, y, z, m, seg_a + seg_ptr[k],
                                seg_b + seg_ptr[k], seg_ptr[k + 1] - seg_ptr[k],
                                bufs, &bcap);
      for (int64_t i = n->offset; i < n->offset + n->count; ++i) {
        float vx = 0, vy = 0, vz = 0;
        for (int64_t j = 0; j < len; ++j) {
          float dx = bufs[0][j] - x[i];
          float dy = bufs[1][j] - y[i];
          float dz = bufs[2][j] - z[i];
          float dist2 = dx * dx + dy * dy + dz * dz;
          float fl;
          if (law == HSRO_LAW_STEP16) {
            /* the Step16 law as a ForceLaw: the r2 > 0 guard of
             * BGQStep16.c:182 replaces updateq's role for self pairs */
            fl = step16_f(dist2, rsm2, fsrrmax2);
          } else {
            fl = law_f32(law, dist2, p->rsm, p->fsrrmax, interp ? &I : NULL);
          }
          float f_over_r = m[i] * bufs[3][j] * fl;
          float updateq = 1.0f;
          updateq *= (float)(dist2 < fsrrmax2);
          vx += updateq * fcoeff * f_over_r * dx;
          vy += updateq * fcoeff * f_over_r * dy;
          vz += updateq * fcoeff * f_over_r * dz;
        }
        ax[i] = vx;
        ay[i] = vy;
        az[i] = vz;
      }
    }
    for (int k = 0; k < 4; ++k) free(bufs[k]);
  }
  if (interp) interp_free(&I);
}

void hsro_force_step16(const hsro_params *p, const hsro_node *t,
                       const float *x, const float *y, const float *z,
                       const float *m, const int64_t *leaves, int64_t nleaves,
                       const int64_t *seg_ptr, const int64_t *seg_a,
                       const int64_t *seg_b, float *ax, float *ay, float *az) {
  const float fsrrmax2 = p->fsrrmax * p->fsrrmax;
  const float rsm2 = p->rsm * p->rsm;
#pragma omp parallel
  {
    float *bufs[4] = {0, 0, 0, 0};
    int64_t bcap = 0;
#pragma omp for schedule(dynamic, 4)
    for (int64_t k = 0; k < nleaves; ++k) {
      const hsro_node *n = &t[leaves[k]];
      int64_t len = gather_list(p, t, x, y, z, m, seg_a + seg_ptr[k],
                                seg_b + seg_ptr[k], seg_ptr[k + 1] - seg_ptr[k],
                                bufs, &bcap);
      for (int64_t i = n->offset; i < n->offset + n->count; ++i) {
        float sx = 0, sy = 0, sz = 0;
        for (int64_t j = 0; j < len; ++j) {
          float dxc = bufs[0][j] - x[i];
          float dyc = bufs[1][j] - y[i];
          float dzc = bufs[2][j] - z[i];
          float r2 = dxc * dxc + dyc * dyc + dzc * dzc;
          float mm = (r2 < fsrrmax2) ? bufs[3][j] : 0.0f;
          float f = (float)(pow((double)(r2 + rsm2), -1.5) -
                            (double)(P5[0] + r2 * (P5[1] + r2 * (P5[2] + r2 *
                            (P5[3] + r2 * (P5[4] + r2 * P5[5]))))));
          f = (r2 > 0.0f) ? mm * f : 0.0f;
          sx = sx + f * dxc;
          sy = sy + f * dyc;
          sz = sz + f * dzc;
        }
        ax[i] = sx;
        ay[i] = sy;
        az[i] = sz;
      }
    }
    for (int k = 0; k < 4; ++k) free(bufs[k]);
  }
}

static double law_f64(double r2, double rsm2, double rc2) {
  if (!(r2 > 0.0) || !(r2 < rc2)) return 0.0;
  double s = r2 + rsm2;
  double poly = (double)P5[0] + r2 * ((double)P5[1] + r2 * ((double)P5[2] +
                r2 * ((double)P5[3] + r2 * ((double)P5[4] + r2 * (double)P5[5]))));
  return 1.0 / (s * sqrt(s)) - poly;
}

void hsro_force_f64(const hsro_params *p, const hsro_node *t,
                    const float *x, const float *y, const float *z,
                    const float *m, const int64_t *leaves, int64_t nleaves,
                    const int64_t *seg_ptr, const int64_t *seg_a,
                    const int64_t *seg_b, double *ax, double *ay, double *az,
                    double *leaf_mom) {
  const double rc2 = (double)(p->fsrrmax * p->fsrrmax);
  const double rsm2 = (double)p->rsm * (double)p->rsm;
#pragma omp parallel
  {
    float *bufs[4] = {0, 0, 0, 0};
    int64_t bcap = 0;
#pragma omp for schedule(dynamic, 4)
    for (int64_t k = 0; k < nleaves; ++k) {
      const hsro_node *n = &t[leaves[k]];
      int64_t len = gather_list(p, t, x, y, z, m, seg_a + seg_ptr[k],
                                seg_b + seg_ptr[k], seg_ptr[k + 1] - seg_ptr[k],
                                bufs, &bcap);
      double mx = 0, my = 0, mz = 0;
      for (int64_t i = n->offset; i < n->offset + n->count; ++i) {
        double sx = 0, sy = 0, sz = 0;
        for (int64_t j = 0; j < len; ++j) {
          double dx = (double)bufs[0][j] - (double)x[i];
          double dy = (double)bufs[1][j] - (double)y[i];
          double dz = (double)bufs[2][j] - (double)z[i];
          double r2 = dx * dx + dy * dy + dz * dz;
          double f = (double)bufs[3][j] * law_f64(r2, rsm2, rc2);
          sx += f * dx;
          sy += f * dy;
          sz += f * dz;
        }
        ax[i] = sx;
        ay[i] = sy;
        az[i] = sz;
        mx += (double)m[i] * sx;
        my += (double)m[i] * sy;
        mz += (double)m[i] * sz;
      }
      if (leaf_mom) {
        leaf_mom[3 * k + 0] = mx;
        leaf_mom[3 * k + 1] = my;
        leaf_mom[3 * k + 2] = mz;
      }
    }
    for (int k = 0; k < 4; ++k) free(bufs[k]);
  }
}

void hsro_direct_f64(int64_t n, const float *x, const float *y, const float *z,
                     const float *m, const int64_t *targets, int64_t ntargets,
                     float rsm, float fsrrmax, double *ax, double *ay,
                     double *az) {
  const double rc2 = (double)(fsrrmax * fsrrmax);
  const double rsm2 = (double)rsm * (double)rsm;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t k = 0; k < ntargets; ++k) {
    int64_t i = targets[k];
    double sx = 0, sy = 0, sz = 0;
    for (int64_t j = 0; j < n; ++j) {
      double dx = (double)x[j] - (double)x[i];
      double dy = (double)y[j] - (double)y[i];
      double dz = (double)z[j] - (double)z[i];
      double r2 = dx * dx + dy * dy + dz * dz;
      double f = (double)m[j] * law_f64(r2, rsm2, rc2);
      sx += f * dx;
      sy += f * dy;
      sz += f * dz;
    }
    ax[k] = sx;
    ay[k] = sy;
    az[k] = sz;
  }
}

/* ------------------------------------------------------------ cell sort */
/* Particles.cxx:402-488: stable counting sort by the unit cell
 * (floorf(x), floorf(y), floorf(z)); index (xx*ng1 + yy)*ng2 + zz; cells
 * outside [0, ng) go to a tail bucket Ng. */
int64_t hsro_cell_sort_perm(int64_t n, const float *x, const float *y,
                            const float *z, const int32_t ng[3],
                            int64_t *perm) {
  int64_t Ng = (int64_t)ng[0] * ng[1] * ng[2];
  int64_t *cnt = (int64_t *)calloc((size_t)(Ng + 1), sizeof(int64_t));
  int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    int xx = (int)floorf(x[i]), yy = (int)floorf(y[i]), zz = (int)floorf(z[i]);
    int64_t c;
    if (xx >= 0 && xx < ng[0] && yy >= 0 && yy < ng[1] && zz >= 0 &&
        zz < ng[2])
      c = ((int64_t)xx * ng[1] + yy) * ng[2] + zz;
    else
      c = Ng;
    ind[i] = c;
    cnt[c]++;
  }
  int64_t inside = n - cnt[Ng];
  int64_t cum = n;
  for (int64_t i = Ng; i >= 0; --i) {
    int64_t tmp = cnt[i];
    cnt[i] = cum - tmp;
    cum -= tmp;
  }
  for (int64_t i = 0; i < n; ++i) perm[i] = cnt[ind[i]]++;
  free(cnt);
  free(ind);
  return inside;
}

/* ------------------------------------------------------------ map1/map2 */
/* Particles.cxx:746-747: pf = powf(pp, 1.0 + 1.0/m_alpha) (double exponent
 * passed to powf as float); prefactor = 1.0 / (m_alpha * adot * pf) with a
 * float product and a double quotient stored as float. */
float hsro_map1_factor(float pp, float tau, float adot, float alpha) {
  float pf = powf(pp, (float)(1.0 + 1.0 / (double)alpha));
  float prefactor = (float)(1.0 / (double)(alpha * adot * pf));
  return prefactor * tau;
}

/* Particles.cxx:753-757: x[i] = x[i] + prefactor * tau * vx[i]. */
void hsro_drift(int64_t n, float *x, float *y, float *z, const float *vx,
                const float *vy, const float *vz, float factor) {
  for (int64_t i = 0; i < n; ++i) {
    x[i] = x[i] + factor * vx[i];
    y[i] = y[i] + factor * vy[i];
    z[i] = z[i] + factor * vz[i];
  }
}

/* Particles.cxx:1230-1233: divscal = gpscal^3 and pi = 4.0*atanf(1.0) stored
 * as float; c = divscal/4.0/pi*fscal*tau*stepFraction in double, stored as
 * float. */
float hsro_map2_coeff(float gpscal, double fscal, double tau,
                      double step_fraction) {
  float divscal = gpscal * gpscal * gpscal;
  float pi = (float)(4.0 * (double)atanf(1.0f));
  return (float)((double)divscal / 4.0 / (double)pi * fscal * tau *
                 step_fraction);
}

/* ------------------------------------------------------------------ CIC */
/* Particles.cxx:371-395 with lo = 0, hi = ng: the cell's row-major index,
 * or `safe` outside the grid. */
static int64_t cic_index(int ix, int iy, int iz, const int32_t *ng, int64_t safe) {
  if (ix >= 0 && ix < ng[0] && iy >= 0 && iy < ng[1] && iz >= 0 && iz < ng[2])
    return ((int64_t)ix * ng[1] + iy) * ng[2] + iz;
  return safe;
}

/* The eight corner cells and weights of Particles.cxx:617-640 / 685-711.
 * ab = 1.0 + (ix - xx): the float difference, the sum in double, stored as
 * float.  Corner q's weight is returned as the reference's expression
 * evaluates it with lead factor s (c for the deposit, the grid value for the
 * interpolation): corner 0 in float, the others in double through the
 * (1.0 - w) factors; *is_float tells which. */
typedef struct {
  int64_t cell[8];
} cic_cells;

static void cic_setup(float xx, float yy, float zz, const int32_t *ng, int64_t safe,
                      cic_cells *cc, float *ab, float *de, float *gh) {
  int ix = (int)floorf(xx), iy = (int)floorf(yy), iz = (int)floorf(zz);
  int ip = ix + 1, jp = iy + 1, kp = iz + 1;
  *ab = (float)(1.0 + (double)((float)ix - xx));
  *de = (float)(1.0 + (double)((float)iy - yy));
  *gh = (float)(1.0 + (double)((float)iz - zz));
  cc->cell[0] = cic_index(ix, iy, iz, ng, safe);
  cc->cell[1] = cic_index(ix, jp, iz, ng, safe);
  cc->cell[2] = cic_index(ix, jp, kp, ng, safe);
  cc->cell[3] = cic_index(ix, iy, kp, ng, safe);
  cc->cell[4] = cic_index(ip, iy, kp, ng, safe);
  cc->cell[5] = cic_index(ip, jp, kp, ng, safe);
  cc->cell[6] = cic_index(ip, jp, iz, ng, safe);
  cc->cell[7] = cic_index(ip, iy, iz, ng, safe);
}

/* s*ab*de*gh and its seven siblings, each as C evaluates the reference's
 * expression left to right (float until the first (1.0 - w) factor, double
 * from there). */
static double cic_term(int q, float s, float ab, float de, float gh) {
  switch (q) {
  case 0: return (double)(s * ab * de * gh);
  case 1: return (double)(s * ab) * (1.0 - de) * gh;
  case 2: return (double)(s * ab) * (1.0 - de) * (1.0 - gh);
  case 3: return (double)(s * ab * de) * (1.0 - gh);
  case 4: return (double)s * (1.0 - ab) * de * (1.0 - gh);
  case 5: return (double)s * (1.0 - ab) * (1.0 - de) * (1.0 - gh);
  case 6: return (double)s * (1.0 - ab) * (1.0 - de) * gh;
  default: return (double)s * (1.0 - ab) * de * gh;
  }
}

/* Particles::cic (Particles.cxx:589-641): rho[Ng + 1] zeroed, then every
 * particle adds c = gpscal^3 times its eight weights in order; corner 0 is a
 * float addition, the others a double addition rounded to float. */
void hsro_cic(int64_t np, const float *x, const float *y, const float *z,
              const int32_t ng[3], float gpscal, float *rho) {
  const int64_t Ng = (int64_t)ng[0] * ng[1] * ng[2];
  const float c = gpscal * gpscal * gpscal;
  memset(rho, 0, sizeof(float) * (size_t)(Ng + 1));
  for (int64_t nn = 0; nn < np; ++nn) {
    cic_cells cc;
    float ab, de, gh;
    cic_setup(x[nn], y[nn], z[nn], ng, Ng, &cc, &ab, &de, &gh);
    for (int q = 0; q < 8; ++q) {
      double t = cic_term(q, c, ab, de, gh);
      float *r = &rho[cc.cell[q]];
      if (q == 0) *r = *r + (float)t;
      else *r = (float)((double)*r + t);
    }
  }
}

/* Particles::inverse_cic (Particles.cxx:645-714) for one component:
 * f = sum of the eight grid values times their weights (same arithmetic as
 * the deposit, grid[Ng] read as 0), then v[nn] += f * fscal * tau in float. */
void hsro_inverse_cic(int64_t np, const float *x, const float *y, const float *z,
                      const int32_t ng[3], const float *grid, float tau, float fscal,
                      float *v) {
  const int64_t Ng = (int64_t)ng[0] * ng[1] * ng[2];
  for (int64_t nn = 0; nn < np; ++nn) {
    cic_cells cc;
    float ab, de, gh;
    cic_setup(x[nn], y[nn], z[nn], ng, Ng, &cc, &ab, &de, &gh);
    float f = 0;
    for (int q = 0; q < 8; ++q) {
      float g = cc.cell[q] == Ng ? 0.0f : grid[cc.cell[q]];
      double t = cic_term(q, g, ab, de, gh);
      if (q == 0) f = f + (float)t;
      else f = (float)((double)f + t);
    }
    v[nn] += f * fscal * tau;
  }
}

/* -------------------------------------------------------------- adjoint */
/* g(r2) and g'(r2) of the Step16 law.  With rsm > 0 the pair term
 * m g(r2 + rsm2) d is smooth through d = 0, so a coincident pair (r2 == 0,
 * whose force term the r2 > 0 guard of BGQStep16.c:180 zeroes) still has
 * the derivative m g(rsm2) I; the guard only removes the pair when rsm = 0. */
static void law_f64_g(double r2, double rsm2, double rc2, double *g,
                      double *gp) {
  if (!(r2 < rc2) || (!(rsm2 > 0.0) && !(r2 > 0.0))) {
    *g = 0.0;
    *gp = 0.0;
    return;
  }
  double s = r2 + rsm2;
  double c[6];
  for (int k = 0; k < 6; ++k) c[k] = (double)P5[k];
  double poly = c[0] + r2 * (c[1] + r2 * (c[2] + r2 * (c[3] + r2 * (c[4] + r2 * c[5]))));
  double dpoly = c[1] + r2 * (2 * c[2] + r2 * (3 * c[3] + r2 * (4 * c[4] + r2 * 5 * c[5])));
  double rs = 1.0 / sqrt(s);
  *g = rs * rs * rs - poly;
  *gp = -1.5 * rs * rs * rs * rs * rs - dpoly;
}

/* FP64 centre of mass of a node's members */
static void node_com(const hsro_node *n, const double *x, const double *y,
                     const double *z, const double *m, double *c, double *M) {
  double sx = 0, sy = 0, sz = 0, sm = 0;
  for (int64_t k = n->offset; k < n->offset + n->count; ++k) {
    sx += m[k] * x[k];
    sy += m[k] * y[k];
    sz += m[k] * z[k];
    sm += m[k];
  }
  c[0] = sx / sm;
  c[1] = sy / sm;
  c[2] = sz / sm;
  *M = sm;
}

double hsro_objective_f64(const hsro_params *p, const hsro_node *t,
                          const double *x, const double *y, const double *z,
                          const double *m, const int64_t *leaves,
                          int64_t nleaves, const int64_t *seg_ptr,
                          const int64_t *seg_a, const int64_t *seg_b,
                          const double *lx, const double *ly,
                          const double *lz) {
  if (tdpts_of(p) != 1) return NAN; /* the adjoint covers the monopole tree */
  const double rc2 = (double)(p->fsrrmax * p->fsrrmax);
  const double rsm2 = (double)p->rsm * (double)p->rsm;
  double J = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : J)
  for (int64_t k = 0; k < nleaves; ++k) {
    const hsro_node *L = &t[leaves[k]];
    for (int64_t s = seg_ptr[k]; s < seg_ptr[k + 1]; ++s) {
      int64_t j0, j1;
      double pc[3], pm = 0;
      int pseudo = seg_b[s] < 0;
      if (pseudo) {
        double M;
        node_com(&t[seg_a[s]], x, y, z, m, pc, &M);
        pm = (double)t[seg_a[s]].ppm[0];
        j0 = 0;
        j1 = 1;
      } else {
        j0 = seg_a[s];
        j1 = seg_a[s] + seg_b[s];
      }
      for (int64_t j = j0; j < j1; ++j) {
        double xj = pseudo ? pc[0] : x[j], yj = pseudo ? pc[1] : y[j],
               zj = pseudo ? pc[2] : z[j], mj = pseudo ? pm : m[j];
        for (int64_t i = L->offset; i < L->offset + L->count; ++i) {
          double dx = xj - x[i], dy = yj - y[i], dz = zj - z[i];
          double g, gp;
          law_f64_g(dx * dx + dy * dy + dz * dz, rsm2, rc2, &g, &gp);
          J += mj * g * (lx[i] * dx + ly[i] * dy + lz[i] * dz);
        }
      }
    }
  }
  return J;
}

void hsro_adjoint_f64(const hsro_params *p, const hsro_node *t,
                      const double *x, const double *y, const double *z,
                      const double *m, const int64_t *leaves, int64_t nleaves,
                      const int64_t *seg_ptr, const int64_t *seg_a,
                      const int64_t *seg_b, const double *lx, const double *ly,
                      const double *lz, double *gx, double *gy, double *gz) {
  if (tdpts_of(p) != 1) return; /* the adjoint covers the monopole tree */
  const double rc2 = (double)(p->fsrrmax * p->fsrrmax);
  const double rsm2 = (double)p->rsm * (double)p->rsm;
  /* serial on purpose: source-side accumulations cross leaves */
  for (int64_t k = 0; k < nleaves; ++k) {
    const hsro_node *L = &t[leaves[k]];
    for (int64_t s = seg_ptr[k]; s < seg_ptr[k + 1]; ++s) {
      int pseudo = seg_b[s] < 0;
      int64_t j0 = pseudo ? 0 : seg_a[s];
      int64_t j1 = pseudo ? 1 : seg_a[s] + seg_b[s];
      double pc[3], M = 1, pm = 0;
      double G[3] = {0, 0, 0};
      if (pseudo) {
        node_com(&t[seg_a[s]], x, y, z, m, pc, &M);
        pm = (double)t[seg_a[s]].ppm[0];
      }
      for (int64_t j = j0; j < j1; ++j) {
        double xj = pseudo ? pc[0] : x[j], yj = pseudo ? pc[1] : y[j],
               zj = pseudo ? pc[2] : z[j], mj = pseudo ? pm : m[j];
        for (int64_t i = L->offset; i < L->offset + L->count; ++i) {
          double dx = xj - x[i], dy = yj - y[i], dz = zj - z[i];
          double g, gp;
          law_f64_g(dx * dx + dy * dy + dz * dz, rsm2, rc2, &g, &gp);
          if (g == 0.0 && gp == 0.0) continue;
          double dl = lx[i] * dx + ly[i] * dy + lz[i] * dz;
          double cx = mj * (g * lx[i] + 2 * gp * dl * dx);
          double cy = mj * (g * ly[i] + 2 * gp * dl * dy);
          double cz = mj * (g * lz[i] + 2 * gp * dl * dz);
          gx[i] -= cx;
          gy[i] -= cy;
          gz[i] -= cz;
          if (pseudo) {
            G[0] += cx;
            G[1] += cy;
            G[2] += cz;
          } else {
            gx[j] += cx;
            gy[j] += cy;
            gz[j] += cz;
          }
        }
      }
      if (pseudo) {
        const hsro_node *N = &t[seg_a[s]];
        for (int64_t q = N->offset; q < N->offset + N->count; ++q) {
          gx[q] += m[q] / M * G[0];
          gy[q] += m[q] / M * G[1];
          gz[q] += m[q] / M * G[2];
        }
      }
    }
  }
}

void hsro_force_f64d(const hsro_params *p, const hsro_node *t,
                     const double *x, const double *y, const double *z,
                     const double *m, const int64_t *leaves, int64_t nleaves,
                     const int64_t *seg_ptr, const int64_t *seg_a,
                     const int64_t *seg_b, double *ax, double *ay, double *az) {
  if (tdpts_of(p) != 1) return; /* the adjoint covers the monopole tree */
  const double rc2 = (double)(p->fsrrmax * p->fsrrmax);
  const double rsm2 = (double)p->rsm * (double)p->rsm;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t k = 0; k < nleaves; ++k) {
    const hsro_node *L = &t[leaves[k]];
    for (int64_t i = L->offset; i < L->offset + L->count; ++i) ax[i] = ay[i] = az[i] = 0.0;
    for (int64_t s = seg_ptr[k]; s < seg_ptr[k + 1]; ++s) {
      int pseudo = seg_b[s] < 0;
      double pc[3], M, pm = 0;
      if (pseudo) {
        node_com(&t[seg_a[s]], x, y, z, m, pc, &M);
        pm = (double)t[seg_a[s]].ppm[0];
      }
      int64_t j0 = pseudo ? 0 : seg_a[s], j1 = pseudo ? 1 : seg_a[s] + seg_b[s];
      for (int64_t j = j0; j < j1; ++j) {
        double xj = pseudo ? pc[0] : x[j], yj = pseudo ? pc[1] : y[j],
               zj = pseudo ? pc[2] : z[j], mj = pseudo ? pm : m[j];
        for (int64_t i = L->offset; i < L->offset + L->count; ++i) {
          double dx = xj - x[i], dy = yj - y[i], dz = zj - z[i];
          double g, gp;
          law_f64_g(dx * dx + dy * dy + dz * dz, rsm2, rc2, &g, &gp);
          ax[i] += mj * g * dx;
          ay[i] += mj * g * dy;
          az[i] += mj * g * dz;
        }
      }
    }
  }
}


/* ------------------------------------------------------- overload refresh */

/* Domain.cxx:42-54, 65-69 with decomposition 1 and position 0. */
void hsro_domain_init(hsro_domain *d) {
  d->phys2grid_pos = d->ng / d->rL;
  d->grid2phys_pos = d->rL / d->ng;
  d->phys2grid_vel = d->ng / (100.0 * d->rL);
  d->grid2phys_vel = 100.0 * d->rL / d->ng;
  d->ng_overload = (int)ceilf(d->oL_in * d->phys2grid_pos);
  d->oL = d->ng_overload * d->grid2phys_pos;
  d->ng_local_total = d->ng + 2 * d->ng_overload + 1;
}

/* The enum NEIGHBOR of BasicDefinition.h:224-264, by name; a name's parts
 * say which face of which dimension the neighbour lies across. */
static const char *const kNeighborNames[26] = {
    "X0",       "X1",       "Y0",       "Y1",       "Z0",       "Z1",
    "X0_Y0",    "X1_Y1",    "X0_Y1",    "X1_Y0",    "Y0_Z0",    "Y1_Z1",
    "Y0_Z1",    "Y1_Z0",    "Z0_X0",    "Z1_X1",    "Z0_X1",    "Z1_X0",
    "X0_Y0_Z0", "X1_Y1_Z1", "X0_Y0_Z1", "X1_Y1_Z0", "X0_Y1_Z0", "X1_Y0_Z1",
    "X0_Y1_Z1", "X1_Y0_Z0"};

/* side[d] = 0 (low face), 1 (high face) or -1 (dimension not named). */
static void neighbor_sides(int n, int side[3]) {
  side[0] = side[1] = side[2] = -1;
  for (const char *c = kNeighborNames[n]; *c; ++c)
    if (*c == 'X' || *c == 'Y' || *c == 'Z') side[*c - 'X'] = c[1] - '0';
}

int64_t hsro_refresh(const hsro_domain *d, float anow, int64_t n, int64_t cap,
                     float *x, float *y, float *z, float *vx, float *vy, float *vz,
                     float *mass, float *phi, int64_t *id, uint16_t *mask,
                     int64_t *alive) {
  float *pos[3] = {x, y, z}, *vel[3] = {vx, vy, vz};
  const int corner_alive = 0 * d->ng;               /* Domain.cxx:74 */
  const int corner = corner_alive - d->ng_overload;  /* Domain.cxx:75 */
  /* Particles::coords_local2global (Particles.cxx:798-830) */
  float pre_vel = d->grid2phys_vel / anow / anow;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      pos[k][i] += corner;
      pos[k][i] *= d->grid2phys_pos;
      vel[k][i] *= pre_vel;
    }
  /* copyAliveIntoVectors (Particles.cxx:959-1030): keep [ca, ca + rla) */
  float ca = corner_alive * d->grid2phys_pos;
  float rla = d->rL / 1;
  float lo = ca, hi = ca + rla;
  int64_t na = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int aa = (x[i] >= lo) * (x[i] < hi) * (y[i] >= lo) * (y[i] < hi) * (z[i] >= lo) *
                   (z[i] < hi);
    if (!aa) continue;
    x[na] = x[i]; y[na] = y[i]; z[na] = z[i];
    vx[na] = vx[i]; vy[na] = vy[i]; vz[na] = vz[i];
    phi[na] = phi[i]; id[na] = id[i]; mask[na] = mask[i];
    mass[na] = 1.0f; /* MC3Extras::refreshParticles pushes 1.0 per particle */
    ++na;
  }
  *alive = na;
  /* ParticleExchange::setParameters / initialize (ParticleExchange.cxx:103,
   * 243-274) with boxSize rL and deadSize oL */
  float box = d->rL, dead = d->oL;
  float step = box / 1;
  float minShare = 0 * step, maxShare = minShare + step;
  if (maxShare > box) maxShare = box;
  float minMine = minShare + dead, maxMine = maxShare - dead;
  /* calculateExchangeRegions / calculateOffsetFactor: a named low face
   * takes [minShare, minMine] and is shifted by +box, a named high face
   * [maxMine, maxShare] and -box, an unnamed dimension the whole share */
  float rmin[26][3], rmax[26][3], off[26][3];
  for (int nb = 0; nb < 26; ++nb) {
    int side[3];
    neighbor_sides(nb, side);
    for (int k = 0; k < 3; ++k) {
      int factor = side[k] == 0 ? 1 : (side[k] == 1 ? -1 : 0);
      rmin[nb][k] = side[k] == 1 ? maxMine : minShare;
      rmax[nb][k] = side[k] == 0 ? minMine : maxShare;
      off[nb][k] = factor * box;
    }
  }
  /* identifyExchangeParticles (ParticleExchange.cxx:542-571), then the
   * exchanges in pairs (n, n+1), (n+1, n), each appending the particles
   * listed for its send side (ParticleExchange.cxx:620-625, 676-695) */
  int64_t total = na;
  for (int nb = 0; nb < 26; ++nb) {
    for (int64_t i = 0; i < na; ++i) {
      if (x[i] > minMine && x[i] < maxMine && y[i] > minMine && y[i] < maxMine &&
          z[i] > minMine && z[i] < maxMine)
        continue;
      if (!(x[i] >= rmin[nb][0] && x[i] <= rmax[nb][0] && y[i] >= rmin[nb][1] &&
            y[i] <= rmax[nb][1] && z[i] >= rmin[nb][2] && z[i] <= rmax[nb][2]))
        continue;
      if (total >= cap) return -1;
      x[total] = x[i] + off[nb][0];
      y[total] = y[i] + off[nb][1];
      z[total] = z[i] + off[nb][2];
      vx[total] = vx[i]; vy[total] = vy[i]; vz[total] = vz[i];
      mass[total] = mass[i]; phi[total] = phi[i]; id[total] = id[i]; mask[total] = mask[i];
      ++total;
    }
  }
  /* copyAllFromVectors (Particles.cxx:309-364): unit masses, then
   * coords_global2local (Particles.cxx:762-796) */
  float pre2 = d->phys2grid_vel * anow * anow;
  for (int64_t i = 0; i < total; ++i) {
    mass[i] = 1.0f;
    for (int k = 0; k < 3; ++k) {
      pos[k][i] *= d->phys2grid_pos;
      pos[k][i] -= corner;
      vel[k][i] *= pre2;
    }
  }
  return total;
}
