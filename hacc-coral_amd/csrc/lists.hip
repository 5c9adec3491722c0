// lists.hip -- bit-exact per-leaf interaction lists on the GPU.
//
// Restates RCBForceTree<TDPTS>::calcInternodeForces / calcInternodeForce
// (RCBForceTree.cxx:924-1197, useRealOA = false): for every leaf inside the
// force box, a right-child-first LIFO walk from the root that
//   * expands the leaf's ancestors to their children (except the leaf),
//   * accepts a node whose angular size passes the tan(oa) test and drops
//     it if its centre is farther than fsrrmax; an accepted node with
//     count <= TDPTS contributes its particles, else its TDPTS
//     pseudo-particles (TDPTS = 1 or 12),
//   * takes a not-accepted leaf's particles,
//   * pushes the children of a not-accepted internal node whose L-infinity
//     gap to the target leaf is <= fsrrmax,
// and finally appends the leaf's own particles.  One thread walks one
// (leaf, ancestor sibling) subtree (two passes: count, then fill), with its
// LIFO stack in HBM laid out level-major so the threads of a wave touch
// neighbouring words.
//
// Compiled with -ffp-contract=off so dist2, l2, dtt2 and the corner
// diagonals round exactly as the reference's x86 float code does.
#include "hsr_internal.h"
#include "wave.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace {

struct ListParams {
  float fsrrmax, fsrrmax2, tan_oa;
  int tdpts;                 // pseudo-particles per node (1 or 12)
  int64_t stack_depth;
  int64_t nthreads;
  int64_t *overflow;
  // the count pass keeps each walk's first `cache` segments (offset or node
  // id, count or -1) at cache_seg[i * nthreads + walk]; the fill pass copies
  // the walks that fit instead of walking them again
  int cache;
  int2 *cache_seg;
};

// Leaf of box b in the right-first order gets slot leaf_base[b] + leaf_rank.
__global__ void k_leaf_slots(const BNode *nodes, int64_t nb,
                             const int32_t *box_of, const int64_t *leaf_base,
                             const float *fmin, const float *fmax,
                             int32_t *slot_canon, const int64_t *cbase) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nb) return;
  const BNode &n = nodes[t];
  if (n.kind == NODE_SPLIT) return;
  int b = box_of[t];
  const float *lo = fmin + 3 * b, *hi = fmax + 3 * b;
  // RCBForceTree.cxx:1166-1170 (strict open interval on either face)
  bool inside = true;
  for (int i = 0; i < 3; ++i) {
    inside &= (n.xmax[i] < hi[i] && n.xmax[i] > lo[i]) ||
              (n.xmin[i] < hi[i] && n.xmin[i] > lo[i]);
  }
  slot_canon[leaf_base[b] + n.leaf_rank] =
      inside ? (int32_t)(cbase[b] + n.canon) : -1;
}

__global__ void k_flag_valid(const int32_t *slot, int64_t n, int32_t *flag) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) flag[k] = slot[k] >= 0 ? 1 : 0;
}

__global__ void k_compact(const int32_t *slot, const int32_t *flag,
                          const int32_t *pos, int64_t n, int32_t *out,
                          int32_t *out_box, const int32_t *slot_box) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n && flag[k]) {
    out[pos[k]] = slot[k];
    out_box[pos[k]] = slot_box[k];
  }
}

__global__ void k_slot_box(int64_t nboxes, const int64_t *leaf_base,
                           int32_t *slot_box) {
  int64_t b = blockIdx.x;
  if (b >= nboxes) return;
  for (int64_t k = leaf_base[b] + threadIdx.x; k < leaf_base[b + 1];
       k += blockDim.x)
    slot_box[k] = (int32_t)b;
}

// The walk of one leaf is a concatenation of independent walks.  Popping an
// ancestor of the leaf pushes its children except the leaf (left, then
// right), so below the ancestor at depth d - 1 the LIFO finishes the right
// child's subtree before the left one's: the list is the walks of the
// siblings of the leaf's ancestors (the other child at each depth d), right
// siblings in increasing d, then left siblings in decreasing d, then the
// leaf's own particles.  One thread walks one (leaf, d) sibling subtree.
// Returns the sibling at depth d (-1 if none) and whether it is the right
// child.
__device__ __forceinline__ int64_t ancestor_sibling(const hsr_node *__restrict__ t, int64_t root,
                                                    int64_t tl, int64_t toff, int64_t tend, int d,
                                                    bool *right) {
  int64_t a = root;
  int64_t sib = -1;
  *right = true;
  // a leaf that is the root is no ancestor of itself: its walk starts at
  // the root like any other (RCBForceTree.cxx:936-962) and takes the root's
  // particles before the own ones are appended
  if (root == tl) return d == 1 ? root : -1;
  for (int i = 1; i <= d; ++i) {
    if (a == tl) return -1;
    const hsr_node &A = t[a];
    const int64_t cl = A.cl, cr = A.cr;
    const bool in_l = cl > 0 && t[cl].count > 0 && t[cl].offset <= toff &&
                      tend <= t[cl].offset + t[cl].count;
    if (in_l) {
      sib = cr;
      *right = true;
      a = cl;
    } else {
      if (cr <= 0) return -1;
      sib = cl;
      *right = false;
      a = cr;
    }
  }
  if (sib <= 0 || t[sib].count <= 0) return -1;
  return sib;
}

template <bool FILL>
__global__ void __launch_bounds__(64)
k_walk(const hsr_node *__restrict__ t, const int32_t *__restrict__ leaves,
       const int32_t *__restrict__ leaf_box, const int64_t *__restrict__ cbase,
       int64_t nleaves, int ndepth, ListParams lp, int32_t *stack, int64_t *nseg_out,
       int64_t *nent_out, uint8_t *side_out, const int64_t *seg_ptr, const int64_t *slot_seg,
       const int64_t *slot_ent, int64_t *seg_a, int64_t *seg_b, int64_t *seg_e) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= nleaves * ndepth) return;
  const int64_t k = gid / ndepth;
  const int d = (int)(gid - k * ndepth) + 1;
  const int64_t tl = leaves[k];
  const int64_t root = cbase[leaf_box[k]];
  const float fsrrmax = lp.fsrrmax, fsrrmax2 = lp.fsrrmax2;
  const float tanOA = lp.tan_oa;
  const hsr_node &L = t[tl];
  const int64_t toff = L.offset, tend = L.offset + L.count;
  if (FILL && nseg_out[gid] <= lp.cache) {
    // the count pass kept this walk's segments: copy them
    const int64_t ns = nseg_out[gid];
    const int64_t so = seg_ptr[k] + slot_seg[gid];
    int64_t nent = slot_ent[gid];
    for (int64_t i = 0; i < ns; ++i) {
      const int2 s = lp.cache_seg[i * lp.nthreads + gid];
      seg_a[so + i] = s.x;
      seg_b[so + i] = s.y;
      seg_e[so + i] = nent;
      nent += s.y < 0 ? lp.tdpts : s.y;
    }
    return;
  }
  bool right;
  const int64_t sib = ancestor_sibling(t, root, tl, toff, tend, d, &right);
  if (!FILL) side_out[gid] = right ? 1 : 0;
  if (sib < 0) {
    if (!FILL) {
      nseg_out[gid] = 0;
      nent_out[gid] = 0;
    }
    return;
  }
  const float lxc0 = L.xc[0], lxc1 = L.xc[1], lxc2 = L.xc[2];
  float lmin[3] = {L.xmin[0], L.xmin[1], L.xmin[2]};
  float lmax[3] = {L.xmax[0], L.xmax[1], L.xmax[2]};
  int32_t *stk = stack + gid;
  const int64_t sstride = lp.nthreads;
  int64_t sp = 0;
  int64_t nseg = 0, nent = FILL ? slot_ent[gid] : 0;
  const int64_t so = FILL ? seg_ptr[k] + slot_seg[gid] : 0;
  bool dead = false;
  stk[0] = (int32_t)sib;
  sp = 1;
  // the LIFO never holds more than (tree depth + 2) nodes; the guard only
  // protects the stack buffer
#define PUSH(v)                                             \
  do {                                                      \
    if (sp >= lp.stack_depth) {                             \
      atomicAdd((unsigned long long *)lp.overflow, 1ull);   \
      dead = true;                                          \
    } else {                                                \
      stk[sp * sstride] = (int32_t)(v);                     \
      ++sp;                                                 \
    }                                                       \
  } while (0)
#define EMIT(A, B, LEN)                                                        \
  do {                                                                         \
    if (FILL) {                                                                \
      seg_a[so + nseg] = (A);                                                  \
      seg_b[so + nseg] = (B);                                                  \
      seg_e[so + nseg] = nent;                                                 \
    } else if (nseg < lp.cache) {                                              \
      lp.cache_seg[nseg * lp.nthreads + gid] = make_int2((int)(A), (int)(B));  \
    }                                                                          \
    ++nseg;                                                                    \
    nent += (LEN);                                                             \
  } while (0)
  while (sp > 0 && !dead) {
    --sp;
    const int64_t tln = stk[sp * sstride];
    const hsr_node &N = t[tln];
    if (tln < tl) {
      // ancestors of tl are exactly the nodes whose particle range contains
      // tl's (node ranges nest or are disjoint)
      if (N.offset <= toff && tend <= N.offset + N.count) {
        const int64_t cr = N.cr, cl = N.cl;
        if (cl != tl && cl > 0 && t[cl].count > 0) PUSH(cl);
        if (cr != tl && cr > 0 && t[cr].count > 0) PUSH(cr);
        continue;
      }
    }
    float dx = N.xc[0] - lxc0;
    float dy = N.xc[1] - lxc1;
    float dz = N.xc[2] - lxc2;
    float dist2 = dx * dx + dy * dy + dz * dz;
    float sx = N.xmax[0] - N.xmin[0];
    float sy = N.xmax[1] - N.xmin[1];
    float sz = N.xmax[2] - N.xmin[2];
    float sx2 = sx * sx, sy2 = sy * sy, sz2 = sz * sz;
    float syz = (sz2 < sy2) ? sz2 : sy2;
    float l2 = (syz < sx2) ? syz : sx2;
    float dtt2 = dist2 * tanOA * tanOA;
    bool looksBig;
    if (l2 > dtt2) {
      looksBig = true;
    } else {
      looksBig = false;
      for (int i = 0; i < 2 && !looksBig; ++i)
        for (int j = 0; j < 2; ++j) {
          float x1 = (i == 0 ? N.xmin[0] : N.xmax[0]) - lxc0;
          float y1 = (j == 0 ? N.xmin[1] : N.xmax[1]) - lxc1;
          float z1 = N.xmin[2] - lxc2;
          float x2 = (i == 0 ? N.xmax[0] : N.xmin[0]) - lxc0;
          float y2 = (j == 0 ? N.xmax[1] : N.xmin[1]) - lxc1;
          float z2 = N.xmax[2] - lxc2;
          float ddx = x1 - x2, ddy = y1 - y2, ddz = z1 - z2;
          float dh2 = ddx * ddx + ddy * ddy + ddz * ddz;
          if (dh2 > dtt2) {
            looksBig = true;
            break;
          }
        }
    }
    if (!looksBig) {
      if (dist2 > fsrrmax2) continue;
      if (N.count <= lp.tdpts) {
        EMIT(N.offset, N.count, N.count);
        continue;
      }
      EMIT(tln, -1, lp.tdpts);
      continue;
    } else if (N.cr == 0 && N.cl == 0) {
      EMIT(N.offset, N.count, N.count);
      continue;
    }
    const int64_t ch[2] = {N.cl, N.cr};
    for (int q = 0; q < 2; ++q) {
      const int64_t c = ch[q];
      if (c > 0 && t[c].count > 0) {
        const hsr_node &Cn = t[c];
        bool close = true;
        for (int i = 0; i < 3; ++i) {
          float dist = 0;
          if (lmax[i] < Cn.xmin[i]) dist = Cn.xmin[i] - lmax[i];
          else if (lmin[i] > Cn.xmax[i]) dist = lmin[i] - Cn.xmax[i];
          if (dist > fsrrmax) {
            close = false;
            break;
          }
        }
        if (close) PUSH(c);
      }
    }
  }
#undef PUSH
#undef EMIT
  if (!FILL) {
    nseg_out[gid] = nseg;
    nent_out[gid] = nent;
  }
}

// Per leaf: the order of its sibling walks (right siblings by increasing
// depth, left ones by decreasing depth; depths past the leaf's hold empty
// walks), their segment and entry offsets in the leaf's list, and the
// leaf's totals with its own particles as the last segment.
__global__ void k_walk_order(const hsr_node *__restrict__ t, const int32_t *__restrict__ leaves,
                             int64_t nleaves, int ndepth, const int64_t *cnt_seg,
                             const int64_t *cnt_ent, const uint8_t *side, int64_t *slot_seg,
                             int64_t *slot_ent, int64_t *nseg, int64_t *nent) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int64_t base = k * ndepth;
  int64_t so = 0, eo = 0;
  for (int i = 0; i < ndepth; ++i)
    if (side[base + i]) {
      slot_seg[base + i] = so;
      slot_ent[base + i] = eo;
      so += cnt_seg[base + i];
      eo += cnt_ent[base + i];
    }
  for (int i = ndepth - 1; i >= 0; --i)
    if (!side[base + i]) {
      slot_seg[base + i] = so;
      slot_ent[base + i] = eo;
      so += cnt_seg[base + i];
      eo += cnt_ent[base + i];
    }
  nseg[k] = so + 1;  // the leaf's own particles
  nent[k] = eo + t[leaves[k]].count;
}

__global__ void k_walk_own(const hsr_node *__restrict__ t, const int32_t *__restrict__ leaves,
                           int64_t nleaves, const int64_t *seg_ptr, const int64_t *ent_ptr,
                           int64_t *seg_a, int64_t *seg_b, int64_t *seg_e) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const hsr_node &L = t[leaves[k]];
  const int64_t s = seg_ptr[k + 1] - 1;
  seg_a[s] = L.offset;
  seg_b[s] = L.count;
  seg_e[s] = ent_ptr[k + 1] - ent_ptr[k] - L.count;
}

// Expand segments into j indices (one wave per listed leaf): particle
// ranges by all lanes, the T pseudo-particles of node a as indices
// count + T a + j (RCBForceTree.cxx:1051-1060).
__global__ void __launch_bounds__(64)
k_expand(int64_t nleaves, int64_t n, int tdpts, const int64_t *__restrict__ seg_ptr,
         const int64_t *__restrict__ seg_a, const int64_t *__restrict__ seg_b,
         const int64_t *__restrict__ ent_ptr, int32_t *__restrict__ jidx) {
  int64_t k = blockIdx.x;
  if (k >= nleaves) return;
  int l = lane_id();
  int64_t e = ent_ptr[k];
  const int64_t s0 = seg_ptr[k], s1 = seg_ptr[k + 1];
  // 64 segments at a time: lane i reads segment s + i, a wave prefix of
  // their entry counts places them, then they are written one by one
  for (int64_t s = s0; s < s1; s += 64) {
    const bool v = s + l < s1;
    const int64_t a = v ? seg_a[s + l] : 0, b = v ? seg_b[s + l] : 0;
    const int64_t len = !v ? 0 : (b < 0 ? (int64_t)tdpts : b);
    int64_t inc = len;  // inclusive prefix over the lanes
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t y = (int64_t)__shfl_up((long long)inc, o);
      if (l >= o) inc += y;
    }
    const int m = (int)min<int64_t>(64, s1 - s);
    for (int j = 0; j < m; ++j) {
      const int64_t aj = (int64_t)__shfl((long long)a, j), bj = (int64_t)__shfl((long long)b, j);
      const int64_t ej = e + (int64_t)__shfl((long long)(inc - len), j);
      if (bj < 0) {
        if (l < tdpts) jidx[ej + l] = (int32_t)(n + tdpts * aj + l);
      } else {
        for (int64_t i = l; i < bj; i += 64) jidx[ej + i] = (int32_t)(aj + i);
      }
    }
    e += (int64_t)__shfl((long long)inc, 63);
  }
}

__global__ void k_pairs(int64_t nleaves, const hsr_node *t,
                        const int32_t *leaves, const int64_t *ent_ptr,
                        int64_t *pairs_out) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t v = 0;
  if (k < nleaves) v = t[leaves[k]].count * (ent_ptr[k + 1] - ent_ptr[k]);
  v = wave_sum_i64(v);
  if (lane_id() == 0 && v)
    atomicAdd((unsigned long long *)pairs_out, (unsigned long long)v);
}

__global__ void k_zero_i64(int64_t *p, int64_t n) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
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

}  // namespace

// BFS node table -> listed leaves, per-leaf lists (segments + j indices).
hsr_status lists_build_device(hsr_ctx *c, const float *min_force_h,
                              const float *max_force_h) {
  hipStream_t st = c->stream;
  TreeState &T = c->tree;
  ListState &LS = c->lists;
  Scratch &S = c->s;
  const int64_t nb = T.nnodes_bfs, nboxes = T.nboxes, n = T.n;
  BNode *nodes = dptr<BNode>(T.bnodes);

  // per-box leaf counts: root.nleaf (read back the roots)
  std::vector<BNode> roots(nboxes);
  HSR_HIP(hipMemcpyAsync(roots.data(), nodes, sizeof(BNode) * nboxes, hipMemcpyDeviceToHost, st));
  HSR_HIP(stream_wait(st));
  std::vector<int64_t> leaf_base(nboxes + 1, 0);
  for (int64_t b = 0; b < nboxes; ++b) leaf_base[b + 1] = leaf_base[b] + roots[b].nleaf;
  const int64_t nslots = leaf_base[nboxes];

  // device scratch: leaf_base, cbase, force boxes, slot arrays
  size_t need = sizeof(int64_t) * 2 * (nboxes + 1) + sizeof(float) * 6 * nboxes +
                sizeof(int32_t) * 4 * (nslots + 1) + 256;
  HSR_HIP(dbuf_reserve(S.g, need));
  int64_t *d_leaf_base = dptr<int64_t>(S.g);
  int64_t *d_cbase = d_leaf_base + nboxes + 1;
  float *d_fmin = reinterpret_cast<float *>(d_cbase + nboxes + 1);
  float *d_fmax = d_fmin + 3 * nboxes;
  int32_t *slot = reinterpret_cast<int32_t *>(d_fmax + 3 * nboxes);
  int32_t *flag = slot + nslots + 1;
  int32_t *pos = flag + nslots + 1;
  int32_t *slot_box = pos + nslots + 1;
  HSR_HIP(hipMemcpyAsync(d_leaf_base, leaf_base.data(), sizeof(int64_t) * (nboxes + 1), hipMemcpyHostToDevice, st));
  HSR_HIP(hipMemcpyAsync(d_cbase, T.box_root.data(), sizeof(int64_t) * (nboxes + 1), hipMemcpyHostToDevice, st));
  HSR_HIP(hipMemcpyAsync(d_fmin, min_force_h, sizeof(float) * 3 * nboxes, hipMemcpyHostToDevice, st));
  HSR_HIP(hipMemcpyAsync(d_fmax, max_force_h, sizeof(float) * 3 * nboxes, hipMemcpyHostToDevice, st));
  const int32_t *box_of = dptr<int32_t>(S.f);  // from the build

  if (nslots > 0) {
    k_leaf_slots<<<nblk(nb, 256), 256, 0, st>>>(nodes, nb, box_of, d_leaf_base, d_fmin, d_fmax, slot, d_cbase);
    k_slot_box<<<(unsigned)nboxes, 256, 0, st>>>(nboxes, d_leaf_base, slot_box);
    k_flag_valid<<<nblk(nslots, 256), 256, 0, st>>>(slot, nslots, flag);
    if (excl_scan(c, flag, pos, nslots) != HSR_OK) return HSR_ERR_HIP;
  }
  int32_t last_pos = 0, last_flag = 0;
  if (nslots > 0) {
    HSR_HIP(hipMemcpyAsync(&last_pos, pos + nslots - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HSR_HIP(hipMemcpyAsync(&last_flag, flag + nslots - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HSR_HIP(stream_wait(st));
  }
  const int64_t nl = nslots > 0 ? (int64_t)last_pos + last_flag : 0;
  LS.nleaves = nl;
  HSR_HIP(dbuf_reserve(LS.leaves, sizeof(int32_t) * (2 * nl + 2)));
  int32_t *leaves = dptr<int32_t>(LS.leaves);
  int32_t *leaf_box = leaves + nl + 1;
  if (nl > 0)
    k_compact<<<nblk(nslots, 256), 256, 0, st>>>(slot, flag, pos, nslots, leaves, leaf_box, slot_box);

  // ---- pass 1: counts
  HSR_HIP(dbuf_reserve(LS.seg_ptr, sizeof(int64_t) * (nl + 1)));
  HSR_HIP(dbuf_reserve(LS.ent_ptr, sizeof(int64_t) * (nl + 1)));
  HSR_HIP(dbuf_reserve(S.h, sizeof(int64_t) * 2 * (nl + 1) + 64));
  int64_t *nseg = dptr<int64_t>(S.h);
  int64_t *nent = nseg + nl + 1;
  const int64_t depth = T.max_depth + 8;
  // one walk per (leaf, ancestor sibling at depth 1 .. max_depth)
  const int ndepth = (int)std::max<int64_t>(1, T.max_depth);
  const int64_t nwalk = std::max<int64_t>(1, nl) * ndepth;
  HSR_HIP(dbuf_reserve(LS.stack, sizeof(int32_t) * (size_t)nwalk * depth));
  HSR_HIP(dbuf_reserve(LS.walk, (sizeof(int64_t) * 4 + 1) * (size_t)nwalk + 64));
  int64_t *cnt_seg = dptr<int64_t>(LS.walk);
  int64_t *cnt_ent = cnt_seg + nwalk;
  int64_t *slot_seg = cnt_ent + nwalk;
  int64_t *slot_ent = slot_seg + nwalk;
  uint8_t *side = reinterpret_cast<uint8_t *>(slot_ent + nwalk);
  ListParams lp;
  lp.fsrrmax = c->p.fsrrmax;
  lp.fsrrmax2 = c->p.fsrrmax * c->p.fsrrmax;
  lp.tan_oa = c->p.tan_oa;
  lp.tdpts = c->p.tdpts;
  lp.stack_depth = depth;
  lp.nthreads = nwalk;
  // segment cache: up to 128 per walk within 2 GiB (at 256^3, 1.1 GB)
  lp.cache = (int)std::min<int64_t>(128, ((int64_t)1 << 31) / ((int64_t)sizeof(int2) * nwalk));
  HSR_HIP(dbuf_reserve(LS.segcache, sizeof(int2) * (size_t)std::max(lp.cache, 1) * (size_t)nwalk));
  lp.cache_seg = dptr<int2>(LS.segcache);
  HSR_HIP(dbuf_reserve(S.counters, 256));
  lp.overflow = dptr<int64_t>(S.counters) + 20;
  k_zero_i64<<<1, 64, 0, st>>>(lp.overflow, 1);
  const hsr_node *cn = dptr<hsr_node>(T.canon_nodes);
  int32_t *stk = dptr<int32_t>(LS.stack);
  if (nl > 0) {
    k_walk<false><<<nblk(nl * ndepth, 64), 64, 0, st>>>(cn, leaves, leaf_box, d_cbase, nl, ndepth, lp,
                                                          stk, cnt_seg, cnt_ent, side, nullptr, nullptr,
                                                          nullptr, nullptr, nullptr, nullptr);
    k_walk_order<<<nblk(nl, 64), 64, 0, st>>>(cn, leaves, nl, ndepth, cnt_seg, cnt_ent, side,
                                              slot_seg, slot_ent, nseg, nent);
    k_zero_i64<<<1, 64, 0, st>>>(nseg + nl, 1);
    k_zero_i64<<<1, 64, 0, st>>>(nent + nl, 1);
    if (excl_scan(c, nseg, dptr<int64_t>(LS.seg_ptr), nl + 1) != HSR_OK) return HSR_ERR_HIP;
    if (excl_scan(c, nent, dptr<int64_t>(LS.ent_ptr), nl + 1) != HSR_OK) return HSR_ERR_HIP;
  }
  int64_t tot[2] = {0, 0};
  if (nl > 0) {
    HSR_HIP(hipMemcpyAsync(&tot[0], dptr<int64_t>(LS.seg_ptr) + nl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HSR_HIP(hipMemcpyAsync(&tot[1], dptr<int64_t>(LS.ent_ptr) + nl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HSR_HIP(stream_wait(st));
  }
  LS.nsegs = tot[0];
  LS.nentries = tot[1];
  // ---- pass 2: fill
  HSR_HIP(dbuf_reserve(LS.seg, sizeof(int64_t) * 3 * (LS.nsegs + 1)));
  int64_t *seg_a = dptr<int64_t>(LS.seg);
  int64_t *seg_b = seg_a + LS.nsegs + 1;
  int64_t *seg_e = seg_b + LS.nsegs + 1;
  HSR_HIP(dbuf_reserve(LS.jidx, sizeof(int32_t) * (LS.nentries + 1)));
  if (nl > 0) {
    k_walk<true><<<nblk(nl * ndepth, 64), 64, 0, st>>>(cn, leaves, leaf_box, d_cbase, nl, ndepth, lp,
                                                         stk, cnt_seg, nullptr, nullptr,
                                                         dptr<int64_t>(LS.seg_ptr),
                                                         slot_seg, slot_ent, seg_a, seg_b, seg_e);
    k_walk_own<<<nblk(nl, 64), 64, 0, st>>>(cn, leaves, nl, dptr<int64_t>(LS.seg_ptr),
                                            dptr<int64_t>(LS.ent_ptr), seg_a, seg_b, seg_e);
    k_expand<<<(unsigned)nl, 64, 0, st>>>(nl, n, c->p.tdpts, dptr<int64_t>(LS.seg_ptr), seg_a, seg_b,
                                          dptr<int64_t>(LS.ent_ptr), dptr<int32_t>(LS.jidx));
  }
  // pairs = sum count * list length
  int64_t *d_pairs = dptr<int64_t>(S.counters) + 16;
  k_zero_i64<<<1, 64, 0, st>>>(d_pairs, 1);
  if (nl > 0) k_pairs<<<nblk(nl, 64), 64, 0, st>>>(nl, cn, leaves, dptr<int64_t>(LS.ent_ptr), d_pairs);
  int64_t ovf = 0;
  HSR_HIP(hipMemcpyAsync(&LS.pairs, d_pairs, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipMemcpyAsync(&ovf, lp.overflow, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HSR_HIP(stream_wait(st));
  HSR_HIP(hipGetLastError());
  if (ovf) {
    hsr_set_error("interaction-list walk exceeded its stack bound");
    return HSR_ERR_STATE;
  }
  return HSR_OK;
}
