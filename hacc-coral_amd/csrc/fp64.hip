// fp64.hip -- FP64 evaluations on the device: an oracle of the leaf kernel
// and the per-leaf momentum sums of the stability check.  Neither is on the
// production path; both walk the same lists as the float kernel.
//
//   hsr_forces_f64     a_i = sum_j m_j f(r2) d in double: float positions and
//                      masses promoted, pseudo-particles at their float
//                      centre with mass ppm, f = (r2 + rsm^2)^-3/2 - P5(r2)
//                      with P5 the float coefficients promoted
//                      (BGQStep16.c:167-187, r2 > 0 guard and r2 < fsrrmax^2
//                      mask as there); optionally sum_i m_i a_i per leaf.
//   hsr_leaf_momentum  sum_i m_i a_i per listed leaf of float accelerations,
//                      accumulated in double in lane order.
#include "hsr_internal.h"
#include "wave.h"

namespace {

constexpr int TILE64 = 512;

struct Law64 {
  double c[7];
  int order;
  double rsm2, rc2;
  int guard;
};

// Fixed-order FP64 wave sum (lane 0 receives the total).
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_down(v, o);
  return v;
}

// One workgroup (one wave) per listed leaf; list entries staged through LDS
// and broadcast to the lanes, which hold the leaf's particles.
__global__ void __launch_bounds__(64)
k_force_f64(int64_t nl, const int32_t *__restrict__ leaves,
            const int64_t *__restrict__ ent_ptr, const int32_t *__restrict__ jidx,
            const hsr_node *__restrict__ cn, const PJ *__restrict__ pj, Law64 L,
            double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
            double *__restrict__ mom) {
  __shared__ float4 tile[TILE64];
  const int64_t k = blockIdx.x;
  if (k >= nl) return;
  const hsr_node &nd = cn[leaves[k]];
  const int64_t e0 = ent_ptr[k], e1 = ent_ptr[k + 1];
  const int l = threadIdx.x;
  double mx = 0, my = 0, mz = 0;
  for (int64_t ib = 0; ib < nd.count; ib += 64) {
    const bool v = ib + l < nd.count;
    const int64_t i = nd.offset + ib + (v ? l : 0);
    const PJ P = pj[i];
    const double xi = P.x, yi = P.y, zi = P.z;
    double sx = 0, sy = 0, sz = 0;
    for (int64_t t0 = e0; t0 < e1; t0 += TILE64) {
      const int nt = (int)min<int64_t>(TILE64, e1 - t0);
      __syncthreads();
      for (int q = l; q < nt; q += 64) {
        const PJ J = pj[jidx[t0 + q]];
        tile[q] = make_float4(J.x, J.y, J.z, J.m);
      }
      __syncthreads();
      for (int q = 0; q < nt; ++q) {
        const float4 a = tile[q];
        const double dx = (double)a.x - xi, dy = (double)a.y - yi, dz = (double)a.z - zi;
        const double r2 = dx * dx + dy * dy + dz * dz;
        double f = 0.0;
        if (r2 < L.rc2 && (r2 > 0.0 || !L.guard)) {
          const double s = r2 + L.rsm2;
          double poly = 0.0;
          for (int c = L.order; c >= 0; --c) poly = poly * r2 + L.c[c];
          f = (double)a.w * (1.0 / (s * sqrt(s)) - poly);
        }
        sx += f * dx;
        sy += f * dy;
        sz += f * dz;
      }
    }
    if (v) {
      ax[i] = sx;
      ay[i] = sy;
      az[i] = sz;
    }
    const double mi = v ? (double)P.m : 0.0;
    mx += wave_sum_f64(mi * sx);
    my += wave_sum_f64(mi * sy);
    mz += wave_sum_f64(mi * sz);
  }
  if (mom && l == 0) {
    mom[3 * k + 0] = mx;
    mom[3 * k + 1] = my;
    mom[3 * k + 2] = mz;
  }
}

__global__ void __launch_bounds__(64)
k_leaf_mom(int64_t nl, const int32_t *__restrict__ leaves, const hsr_node *__restrict__ cn,
           const PJ *__restrict__ pj, const float *__restrict__ ax,
           const float *__restrict__ ay, const float *__restrict__ az,
           double *__restrict__ mom) {
  const int64_t k = blockIdx.x;
  if (k >= nl) return;
  const hsr_node &nd = cn[leaves[k]];
  const int l = threadIdx.x;
  double mx = 0, my = 0, mz = 0;
  for (int64_t ib = 0; ib < nd.count; ib += 64) {
    const bool v = ib + l < nd.count;
    const int64_t i = nd.offset + ib + (v ? l : 0);
    const double m = v ? (double)pj[i].m : 0.0;
    mx += wave_sum_f64(m * (double)(v ? ax[i] : 0.f));
    my += wave_sum_f64(m * (double)(v ? ay[i] : 0.f));
    mz += wave_sum_f64(m * (double)(v ? az[i] : 0.f));
  }
  if (l == 0) {
    mom[3 * k + 0] = mx;
    mom[3 * k + 1] = my;
    mom[3 * k + 2] = mz;
  }
}

__global__ void k_fill_f64(int64_t n, double *a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0.0;
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// ----------------------------------------------------------- FP64 at positions
// The FP64 oracle of the differentiation checks: forces and their
// vector-Jacobian product at given double positions with the context's tree
// and lists frozen.  Sources are records PD[n + nnodes] (x, y, z, m) in the
// layout of PJ: particles at the given positions with their masses, node k's
// pseudo-particle at the FP64 centre of mass of its members' given positions
// with mass ppm (RCBForceTree.cxx:525-569 with the centre evaluated in FP64).
struct PD {
  double x, y, z, m;
};

struct LawAt {
  double c[7];  // P(r2), a0..a6
  int order;
  double rsm2, rc2;
  int guard;    // rsm = 0: pairs at r2 = 0 are skipped
};

// g = s^-3/2 - P(r2) and g' = dg/dr2 at r2 (zero beyond the cutoff).
__device__ __forceinline__ void law_at(const LawAt &L, double r2, double &g, double &gp) {
  if (!(r2 < L.rc2) || (L.guard && !(r2 > 0.0))) {
    g = gp = 0.0;
    return;
  }
  const double s = r2 + L.rsm2;
  double poly = 0.0, dpoly = 0.0;
  for (int c = L.order; c >= 0; --c) poly = poly * r2 + L.c[c];
  for (int c = L.order; c >= 1; --c) dpoly = dpoly * r2 + c * L.c[c];
  const double rs = 1.0 / sqrt(s);
  const double r3 = rs * rs * rs;
  g = r3 - poly;
  gp = -1.5 * r3 * rs * rs - dpoly;
}

__global__ void k_pd_particles(int64_t n, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ z,
                               const PJ *__restrict__ pj, PD *__restrict__ pd) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) pd[i] = PD{x[i], y[i], z[i], (double)pj[i].m};
}

// One wave per node: the FP64 member sums in a fixed order (lanes strided,
// then an xor tree); mnode receives the member mass.
__global__ void __launch_bounds__(64)
k_pd_nodes(int64_t nn, const hsr_node *__restrict__ cn, PD *__restrict__ pd, int64_t n,
           double *__restrict__ mnode) {
  const int64_t k = blockIdx.x;
  if (k >= nn) return;
  const hsr_node &nd = cn[k];
  const int l = threadIdx.x;
  double sx = 0, sy = 0, sz = 0, sm = 0;
  for (int64_t i = nd.offset + l; i < nd.offset + nd.count; i += 64) {
    const PD q = pd[i];
    sx += q.m * q.x;
    sy += q.m * q.y;
    sz += q.m * q.z;
    sm += q.m;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sx += __shfl_xor(sx, o);
    sy += __shfl_xor(sy, o);
    sz += __shfl_xor(sz, o);
    sm += __shfl_xor(sm, o);
  }
  if (l == 0) {
    const bool ok = nd.count > 0 && sm > 0.0;
    pd[n + k] = PD{ok ? sx / sm : 0.0, ok ? sy / sm : 0.0, ok ? sz / sm : 0.0, (double)nd.ppm[0]};
    mnode[k] = sm;
  }
}

// Forces: one wave per listed leaf, 64 targets at a time, every list entry
// in list order (summation order fixed).
__global__ void __launch_bounds__(64)
k_force_f64_at(int64_t nl, const int32_t *__restrict__ leaves, const int64_t *__restrict__ ent_ptr,
               const int32_t *__restrict__ jidx, const hsr_node *__restrict__ cn,
               const PD *__restrict__ pd, LawAt L, double *__restrict__ ax,
               double *__restrict__ ay, double *__restrict__ az) {
  const int64_t k = blockIdx.x;
  if (k >= nl) return;
  const hsr_node &nd = cn[leaves[k]];
  const int64_t e0 = ent_ptr[k], e1 = ent_ptr[k + 1];
  const int l = threadIdx.x;
  for (int64_t ib = 0; ib < nd.count; ib += 64) {
    const bool v = ib + l < nd.count;
    const int64_t i = nd.offset + ib + (v ? l : 0);
    const PD P = pd[i];
    double sx = 0, sy = 0, sz = 0;
    for (int64_t e = e0; e < e1; ++e) {
      const PD J = pd[jidx[e]];
      const double dx = J.x - P.x, dy = J.y - P.y, dz = J.z - P.z;
      double g, gp;
      law_at(L, dx * dx + dy * dy + dz * dz, g, gp);
      sx += J.m * g * dx;
      sy += J.m * g * dy;
      sz += J.m * g * dz;
    }
    if (v) {
      ax[i] = sx;
      ay[i] = sy;
      az[i] = sz;
    }
  }
}

// VJP, dJ/dx for J = sum_i lambda_i . F_i: per pair c = m_j [g lam_i +
// 2 g' (d . lam_i) d], d = x_j - x_i; the target gets -c, the source +c
// (summed over the wave's targets, then added atomically: the order of
// those FP64 additions varies, far below the checks' tolerances); a
// pseudo-particle's sum goes to its node (gs[n + node]).
__global__ void __launch_bounds__(64)
k_adjoint_f64_at(int64_t nl, const int32_t *__restrict__ leaves,
                 const int64_t *__restrict__ ent_ptr, const int32_t *__restrict__ jidx,
                 const hsr_node *__restrict__ cn, const PD *__restrict__ pd, LawAt L,
                 const double *__restrict__ lx, const double *__restrict__ ly,
                 const double *__restrict__ lz, double *__restrict__ gx,
                 double *__restrict__ gy, double *__restrict__ gz) {
  const int64_t k = blockIdx.x;
  if (k >= nl) return;
  const hsr_node &nd = cn[leaves[k]];
  const int64_t e0 = ent_ptr[k], e1 = ent_ptr[k + 1];
  const int l = threadIdx.x;
  for (int64_t ib = 0; ib < nd.count; ib += 64) {
    const bool v = ib + l < nd.count;
    const int64_t i = nd.offset + ib + (v ? l : 0);
    const PD P = pd[i];
    const double la = v ? lx[i] : 0.0, lb = v ? ly[i] : 0.0, lc = v ? lz[i] : 0.0;
    double tx = 0, ty = 0, tz = 0;
    for (int64_t e = e0; e < e1; ++e) {
      const int32_t j = jidx[e];
      const PD J = pd[j];
      const double dx = J.x - P.x, dy = J.y - P.y, dz = J.z - P.z;
      double g, gp;
      law_at(L, dx * dx + dy * dy + dz * dz, g, gp);
      const double dl = la * dx + lb * dy + lc * dz;
      double cx = J.m * (g * la + 2.0 * gp * dl * dx);
      double cy = J.m * (g * lb + 2.0 * gp * dl * dy);
      double cz = J.m * (g * lc + 2.0 * gp * dl * dz);
      if (!v) cx = cy = cz = 0.0;
      tx -= cx;
      ty -= cy;
      tz -= cz;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        cx += __shfl_xor(cx, o);
        cy += __shfl_xor(cy, o);
        cz += __shfl_xor(cz, o);
      }
      if (l == 0 && (cx != 0.0 || cy != 0.0 || cz != 0.0)) {
        atomicAdd(gx + j, cx);
        atomicAdd(gy + j, cy);
        atomicAdd(gz + j, cz);
      }
    }
    if (v) {
      atomicAdd(gx + i, tx);
      atomicAdd(gy + i, ty);
      atomicAdd(gz + i, tz);
    }
  }
}

// Pseudo-particle sums to the members: a member k of node m moves the
// node's centre by m_k / M_m, so g_k += m_k * sum over the nodes holding k
// of G_m / M_m (walking down from the root of a single tree).
__global__ void k_push_f64_at(int64_t n, const hsr_node *__restrict__ cn,
                              const double *__restrict__ mnode, const PD *__restrict__ pd,
                              double *__restrict__ gx, double *__restrict__ gy,
                              double *__restrict__ gz) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double ax = 0, ay = 0, az = 0;
  int64_t node = 0;
  for (int depth = 0; depth < 256; ++depth) {
    const hsr_node &nd = cn[node];
    const double M = mnode[node];
    if (M > 0.0) {
      ax += gx[n + node] / M;
      ay += gy[n + node] / M;
      az += gz[n + node] / M;
    }
    if (nd.cl <= 0 && nd.cr <= 0) break;
    const hsr_node &L = cn[nd.cl];
    node = (nd.cl > 0 && k >= L.offset && k < L.offset + L.count) ? nd.cl : nd.cr;
    if (node <= 0) break;
  }
  const double m = pd[k].m;
  gx[k] += m * ax;
  gy[k] += m * ay;
  gz[k] += m * az;
}

}  // namespace

// Shared set-up of the two FP64-at-positions evaluations: checks, the law,
// and the source records.
static hsr_status f64_at_setup(hsr_ctx *c, const double *x, const double *y, const double *z,
                               LawAt *L, PD **pd, double **mnode) {
  if (!c->listed) {
    hsr_set_error("FP64 at positions: no lists (call hsr_lists first)");
    return HSR_ERR_STATE;
  }
  if (c->law.newton || c->law.kind != LAWK_POLY || c->p.tdpts != 1 || c->tree.nboxes != 1) {
    hsr_set_error("FP64 at positions: covers the polynomial laws on a single monopole tree");
    return HSR_ERR_UNSUPPORTED;
  }
  for (int k = 0; k < 7; ++k) L->c[k] = (double)c->law.c[k];
  L->order = c->law.order;
  L->rsm2 = (double)c->p.rsm * (double)c->p.rsm;
  L->rc2 = (double)c->law.rc2;
  L->guard = L->rsm2 > 0.0 ? 0 : 1;
  const int64_t n = c->tree.n, nn = c->tree.nnodes;
  HSR_HIP(dbuf_reserve(c->f64_at, sizeof(PD) * (size_t)(n + nn + 1) + sizeof(double) * (size_t)(nn + 1)));
  *pd = dptr<PD>(c->f64_at);
  *mnode = reinterpret_cast<double *>(*pd + n + nn + 1);
  hipStream_t st = c->stream;
  if (n) k_pd_particles<<<nblk(n, 256), 256, 0, st>>>(n, x, y, z, dptr<PJ>(c->tree.pj), *pd);
  if (nn) k_pd_nodes<<<(unsigned)nn, 64, 0, st>>>(nn, dptr<hsr_node>(c->tree.canon_nodes), *pd, n, *mnode);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

hsr_status forces_f64_at_run(hsr_ctx *c, const double *x, const double *y, const double *z,
                             double *ax, double *ay, double *az) {
  LawAt L;
  PD *pd;
  double *mnode;
  hsr_status s = f64_at_setup(c, x, y, z, &L, &pd, &mnode);
  if (s) return s;
  hipStream_t st = c->stream;
  const int64_t n = c->tree.n, nl = c->lists.nleaves;
  if (n) {
    k_fill_f64<<<nblk(n, 256), 256, 0, st>>>(n, ax);
    k_fill_f64<<<nblk(n, 256), 256, 0, st>>>(n, ay);
    k_fill_f64<<<nblk(n, 256), 256, 0, st>>>(n, az);
  }
  if (nl)
    k_force_f64_at<<<(unsigned)nl, 64, 0, st>>>(nl, dptr<int32_t>(c->lists.leaves),
                                                dptr<int64_t>(c->lists.ent_ptr), dptr<int32_t>(c->lists.jidx),
                                                dptr<hsr_node>(c->tree.canon_nodes), pd, L, ax, ay, az);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

// g must hold n + nnodes doubles per component (the node part is scratch).
hsr_status adjoint_f64_at_run(hsr_ctx *c, const double *x, const double *y, const double *z,
                              const double *lx, const double *ly, const double *lz, double *gx,
                              double *gy, double *gz) {
  LawAt L;
  PD *pd;
  double *mnode;
  hsr_status s = f64_at_setup(c, x, y, z, &L, &pd, &mnode);
  if (s) return s;
  hipStream_t st = c->stream;
  const int64_t n = c->tree.n, nn = c->tree.nnodes, nl = c->lists.nleaves;
  const int64_t ng = n + nn;
  if (ng) {
    k_fill_f64<<<nblk(ng, 256), 256, 0, st>>>(ng, gx);
    k_fill_f64<<<nblk(ng, 256), 256, 0, st>>>(ng, gy);
    k_fill_f64<<<nblk(ng, 256), 256, 0, st>>>(ng, gz);
  }
  if (nl)
    k_adjoint_f64_at<<<(unsigned)nl, 64, 0, st>>>(nl, dptr<int32_t>(c->lists.leaves),
                                                  dptr<int64_t>(c->lists.ent_ptr), dptr<int32_t>(c->lists.jidx),
                                                  dptr<hsr_node>(c->tree.canon_nodes), pd, L, lx, ly, lz,
                                                  gx, gy, gz);
  if (n)
    k_push_f64_at<<<nblk(n, 256), 256, 0, st>>>(n, dptr<hsr_node>(c->tree.canon_nodes), mnode, pd, gx, gy,
                                               gz);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

hsr_status forces_f64_run(hsr_ctx *c, double *ax, double *ay, double *az, double *mom) {
  hipStream_t st = c->stream;
  const int64_t n = c->tree.n, nl = c->lists.nleaves;
  Law64 L;
  for (int k = 0; k < 7; ++k) L.c[k] = (double)c->law.c[k];
  L.order = c->law.order;
  L.rsm2 = (double)c->p.rsm * (double)c->p.rsm;
  L.rc2 = (double)c->law.rc2;
  L.guard = 1;
  if (c->law.newton || c->law.kind != LAWK_POLY) {
    hsr_set_error("hsr_forces_f64: the FP64 evaluation covers the polynomial laws");
    return HSR_ERR_UNSUPPORTED;
  }
  if (n) {
    k_fill_f64<<<nblk(n, 256), 256, 0, st>>>(n, ax);
    k_fill_f64<<<nblk(n, 256), 256, 0, st>>>(n, ay);
    k_fill_f64<<<nblk(n, 256), 256, 0, st>>>(n, az);
  }
  if (nl)
    k_force_f64<<<(unsigned)nl, 64, 0, st>>>(nl, dptr<int32_t>(c->lists.leaves), dptr<int64_t>(c->lists.ent_ptr),
                                             dptr<int32_t>(c->lists.jidx), dptr<hsr_node>(c->tree.canon_nodes),
                                             dptr<PJ>(c->tree.pj), L, ax, ay, az, mom);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

hsr_status leaf_momentum_run(hsr_ctx *c, const float *ax, const float *ay, const float *az,
                             double *mom) {
  const int64_t nl = c->lists.nleaves;
  if (nl)
    k_leaf_mom<<<(unsigned)nl, 64, 0, c->stream>>>(nl, dptr<int32_t>(c->lists.leaves),
                                                  dptr<hsr_node>(c->tree.canon_nodes), dptr<PJ>(c->tree.pj),
                                                  ax, ay, az, mom);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}
