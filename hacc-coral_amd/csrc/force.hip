// force.hip -- short-range leaf kernel for gfx950 (FP32 VALU; no MFMA: the
// pair force is not a dense contraction).
//
// Pair law (BGQStep16.c:167-187, the 5th-order polynomial of ForceLaw.cxx
// :107-116 at POLY_ORDER 5):
//     f(r2) = m_j * [ (r2 + rsm^2)^(-3/2) - P(r2) ] * [r2 < rc2] (* [r2 > 0])
//     F_i  += f * (x_j - x_i)
// summed over every entry j of leaf i's interaction list (nbody1,
// RCBForceTree.cxx:599-619).
//
// Mapping: one wave per work item (a listed leaf, or a chunk of a long
// list).  The leaf's i-particles sit in registers, R per lane, in G lane
// groups; each group walks its own stride of the j-tile staged in LDS from
// the list's j indices (coalesced 16-byte records).  Group partials are
// reduced in a fixed order, so results are bitwise reproducible.
#include "hsr_internal.h"
#include "wave.h"

#include <algorithm>

namespace {

constexpr int TILE = 512;

template <int ORDER>
__device__ __forceinline__ float poly(const ForceLawDev &L, float r2) {
  float p;
  if (ORDER == 6) {
    p = fmaf(L.c[6], r2, L.c[5]);
    p = fmaf(p, r2, L.c[4]);
  } else {
    p = fmaf(L.c[5], r2, L.c[4]);
  }
  p = fmaf(p, r2, L.c[3]);
  p = fmaf(p, r2, L.c[2]);
  p = fmaf(p, r2, L.c[1]);
  p = fmaf(p, r2, L.c[0]);
  return p;
}

struct Item {
  int32_t leaf;       // listed-leaf index
  int32_t part;       // chunk index within the leaf's list
  int64_t e0, e1;     // list entry range
};

// Work item -> force partials.  R i-particles per lane; lane groups of
// `lpg` lanes split the j-range by stride G.
template <int R, int ORDER, bool GUARD>
__global__ void __launch_bounds__(64)
k_force(const hsr_node *__restrict__ cn, const int32_t *__restrict__ leaves,
        const int64_t *__restrict__ ent_ptr, const int32_t *__restrict__ jidx,
        const PJ *__restrict__ pj, const float *__restrict__ xs,
        const float *__restrict__ ys, const float *__restrict__ zs,
        int64_t nleaves, ForceLawDev L, float *__restrict__ fx,
        float *__restrict__ fy, float *__restrict__ fz) {
  __shared__ float4 tile[TILE];
  __shared__ float red[3][64 * R];
  const int64_t k = blockIdx.x;
  if (k >= nleaves) return;
  const int l = lane_id();
  const hsr_node &nd = cn[leaves[k]];
  const int64_t off = nd.offset, cnt = nd.count;
  const int64_t e0 = ent_ptr[k], e1 = ent_ptr[k + 1];
  const float eps2 = L.eps2, rc2 = L.rc2;

  for (int64_t ib = 0; ib < cnt; ib += 64 * R) {
    const int ni = (int)min<int64_t>(64 * R, cnt - ib);
    // lane groups: lpg lanes each holding R i's cover ni particles
    const int lpg = (ni + R - 1) / R;
    const int G = max(1, 64 / lpg);
    const int g = l / lpg, tq = l - g * lpg;
    const bool active = g < G;
    float xi[R], yi[R], zi[R], ax[R], ay[R], az[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int ii = tq + r * lpg;
      bool v = active && ii < ni;
      xi[r] = v ? xs[off + ib + ii] : 1e30f;
      yi[r] = v ? ys[off + ib + ii] : 1e30f;
      zi[r] = v ? zs[off + ib + ii] : 1e30f;
      ax[r] = ay[r] = az[r] = 0.f;
    }
    for (int64_t t0 = e0; t0 < e1; t0 += TILE) {
      const int nt = (int)min<int64_t>(TILE, e1 - t0);
      __syncthreads();
      for (int q = l; q < nt; q += 64) {
        PJ p = pj[jidx[t0 + q]];
        tile[q] = make_float4(p.x, p.y, p.z, p.m);
      }
      __syncthreads();
      if (active) {
        for (int q = g; q < nt; q += G) {
          const float4 p = tile[q];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            float dx = p.x - xi[r];
            float dy = p.y - yi[r];
            float dz = p.z - zi[r];
            float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            float s = r2 + eps2;
            float y = __builtin_amdgcn_rsqf(s);
            float y2 = y * y;
            float f = fmaf(y2, y, -poly<ORDER>(L, r2));
            float mm = (r2 < rc2) ? p.w : 0.f;
            if (GUARD) mm = (r2 > 0.f) ? mm : 0.f;
            f *= mm;
            ax[r] = fmaf(f, dx, ax[r]);
            ay[r] = fmaf(f, dy, ay[r]);
            az[r] = fmaf(f, dz, az[r]);
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
          int ii = tq + r * lpg;
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
      fx[off + ib + ii] = red[0][ii];
      fy[off + ib + ii] = red[1][ii];
      fz[off + ib + ii] = red[2][ii];
    }
  }
}

__global__ void k_fill3(int64_t n, float *a, float *b, float *c) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = b[k] = c[k] = 0.f;
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

template <int R, int ORDER, bool GUARD>
void launch_force(hsr_ctx *c, float *fx, float *fy, float *fz) {
  ListState &LS = c->lists;
  TreeState &T = c->tree;
  k_force<R, ORDER, GUARD><<<(unsigned)LS.nleaves, 64, 0, c->stream>>>(
      dptr<hsr_node>(T.canon_nodes), dptr<int32_t>(LS.leaves),
      dptr<int64_t>(LS.ent_ptr), dptr<int32_t>(LS.jidx), dptr<PJ>(T.pj),
      dptr<float>(T.xs), dptr<float>(T.ys), dptr<float>(T.zs), LS.nleaves,
      c->law, fx, fy, fz);
}

}  // namespace

hsr_status force_run(hsr_ctx *c, float *fx, float *fy, float *fz) {
  hipStream_t st = c->stream;
  int64_t n = c->tree.n;
  c->kernel_timed = false;
  if (n > 0)
    k_fill3<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, fx, fy, fz);
  if (c->lists.nleaves == 0) return HSR_OK;
  const bool guard = c->law.guard != 0;
  const int order = c->law.order;
  // events bracket the leaf-kernel launch alone (the roofline's duration)
  HSR_HIP(hipEventRecord(c->ev[6], st));
  if (order == 6) {
    if (guard) launch_force<2, 6, true>(c, fx, fy, fz);
    else launch_force<2, 6, false>(c, fx, fy, fz);
  } else {
    if (guard) launch_force<2, 5, true>(c, fx, fy, fz);
    else launch_force<2, 5, false>(c, fx, fy, fz);
  }
  HSR_HIP(hipGetLastError());
  HSR_HIP(hipEventRecord(c->ev[7], st));
  c->kernel_timed = true;
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

hsr_status adjoint_run(hsr_ctx *c, const float *lx, const float *ly,
                       const float *lz, float *gx, float *gy, float *gz) {
  (void)c; (void)lx; (void)ly; (void)lz; (void)gx; (void)gy; (void)gz;
  hsr_set_error("hsr_adjoint: not available in this build");
  return HSR_ERR_UNSUPPORTED;
}
