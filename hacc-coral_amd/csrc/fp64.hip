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

}  // namespace

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
