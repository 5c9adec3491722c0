// hsr_internal.h -- device data layout and kernel launchers shared by the
// translation units of libhaccsr.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/hacc_sr.h"

#define HSR_WAVE 64

// Node of the breadth-first build table.  Nodes of one level are contiguous;
// children of a level are appended as the next level.  The canonical
// (reference) numbering is derived after the build.
struct BNode {
  int64_t count, offset;
  int32_t parent, cl, cr;  // BFS ids, -1 if none
  int32_t kind;            // NODE_LEAF / NODE_SPLIT / NODE_DEGEN
  int32_t split_d;
  int32_t depth;
  int64_t is;              // partition result (left count)
  float xmin[3], xmax[3], xc[3];
  float tdr, ppm;
  int64_t canon;           // canonical id (allocation order of the unthreaded build)
  int64_t sa;              // allocating nodes in the subtree (incl. itself)
  int64_t a;               // allocating nodes before this one in preorder
  int64_t nleaf;           // visited leaves in the subtree
  int64_t leaf_rank;       // rank in calcInternodeForces' right-first order
  int64_t blk0;            // first block of this node in a level block table
  int64_t nblk;
};

enum { NODE_LEAF = 0, NODE_SPLIT = 1, NODE_DEGEN = 2 };

// Exact centre-of-mass accumulators of one block of a big node.
// (a block's node is in a separate int32 table; its element range follows
// from the node's offset, count and first block)
struct CmBlock {
  float approx[4];         // float sums of x, y, z, m addends
  float ein[4];            // approximate exclusive prefix within the node
  uint64_t d0[4], d1[4];   // fixed-grid maps
  int32_t E[4];            // grid exponent each map was built for
  int32_t bad[4];
  float bmin[3], bmax[3];
  // the leader block (first of each 64-block window of a node) also holds the
  // window's composed map, valid when all its blocks share one grid
  uint64_t wd0[4], wd1[4];
  int32_t wE[4];
  int32_t wok[4];
  int32_t lsb[4];          // finest lsb exponent of the addends (INT32_MIN: a negative one)
  uint64_t units[4];       // their sum in units of 2^lsb (~0: not an integer there)
  // the leader block also holds its window's reductions (a chunk of up to
  // 64 blocks) and the exclusive prefix of the approximate sums over the
  // node's chunks
  float capprox[4], cbase[4];
  float cbmin[3], cbmax[3];
  int32_t clsb[4];         // finest lsb over the chunk (INT32_MIN: a negative addend)
  uint64_t cunits[4];      // its sum in units of 2^clsb, saturated at 2^53
};

struct PartBlock {
  int64_t cnt;             // keys < pivot in the block
  int64_t base;            // exclusive prefix within the node
};

// Growable device buffer.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

hipError_t dbuf_reserve(DBuf &b, size_t bytes);
void dbuf_free(DBuf &b);

template <class T>
T *dptr(DBuf &b) {
  return reinterpret_cast<T *>(b.p);
}

// Sorted j-particle record (x, y, z, m) for the leaf kernel; the TDPTS
// pseudo-particles of node k follow the particles at count + TDPTS k + j.
struct __align__(16) PJ {
  float x, y, z, m;
};

// Kinds of grid-force subtraction the leaf kernel evaluates.
enum { LAWK_POLY = 0, LAWK_FIT = 1, LAWK_INTERP = 2 };

struct ForceLawDev {
  float c[7];     // polynomial P(r2) coefficients a0..a6
  int order;      // 5 or 6 (0 = no polynomial)
  float eps2;     // rsm^2
  float rc2;      // strict cutoff: pairs with r2 < rc2 contribute
  int guard;      // 1: r2 > 0 guard (needed when eps2 == 0)
  int newton;     // 1: f = r^-3 evaluated without softening
  int kind;       // LAWK_*: polynomial, FGrid::fgor per pair, or its table
  int n_interp;   // LAWK_INTERP: table size
  // LAWK_INTERP: the caller's table, device bins {r2[i], f[i], f[i+1]-f[i], 0}
  // for i < n_interp - 1, and the range FGridEvalInterp derives from it
  const float4 *bins;
  float t_r2min, t_r2max, t_oodr2;
};

// Tree-build state kept in the context between calls.
struct TreeState {
  int64_t n = 0;
  int64_t nnodes_bfs = 0;
  int64_t nnodes = 0;         // canonical count
  int64_t max_depth = 0;
  std::vector<int64_t> level_start;  // BFS id ranges per level
  DBuf bnodes;                // BNode[nnodes_bfs]
  DBuf canon_nodes;           // hsr_node[nnodes]
  DBuf pfin;                  // int32 final permutation (sorted -> input)
  DBuf pj;                    // PJ[n + tdpts * nnodes]
  DBuf ppm12;                 // float[12 * nnodes]: TDPTS = 12 pseudo masses
  DBuf xs, ys, zs, ms;        // sorted float arrays
  bool unit = false;          // every mass of the last build was exactly 1.0f
  int64_t nboxes = 1;
  std::vector<int64_t> box_root;  // BFS id of each box root
};

struct ListState {
  int64_t nleaves = 0;        // listed leaves
  int64_t nsegs = 0;
  int64_t nentries = 0;
  int64_t pairs = 0;
  DBuf leaves;                // int32 canonical ids of listed leaves (reference order)
  DBuf seg_ptr;               // int64 [nleaves+1]
  DBuf seg;                   // int2-like (int64 a, int64 b) pairs
  DBuf ent_ptr;               // int64 [nleaves+1] list-entry CSR
  DBuf jidx;                  // int32 [nentries] index into PJ
  DBuf items;                 // work items for the force kernel (sorted)
  DBuf item_keys;             // sort scratch
  DBuf plan;                  // per-leaf chunk plan
  DBuf part;                  // chunk partial forces of long lists
  int64_t nitems = 0, nmulti = 0, chunk = 0;
  int64_t plan_stride = 0;    // per-leaf plan arrays' stride (grows only)
  int64_t *multi = nullptr, *counts = nullptr, *leaf_pbase = nullptr;
  int32_t *leaf_nchunk = nullptr;
  DBuf stack;                 // DFS stacks
  DBuf walk;                  // per (leaf, sibling walk): counts and offsets
  DBuf segcache;              // the first segments each walk finds (count pass)
  DBuf iperm;                 // int32 [n]: target order inside each listed leaf
};

// Adjoint plan: item queues of the target and source passes, the
// transposed (source-major) lists and the pseudo-particle segments.
struct AdjState {
  int64_t gen = -1;
  DBuf fw, fw_items, tr, tr_items, seg, tj, par, counts, qctr, part, H, W;
  int64_t n_fw_items = 0, n_fw_leaves = 0, n_tr_items = 0, n_tr_leaves = 0;
  int64_t nsp = 0, npseudo = 0, fw_part = 0, tr_part = 0;
  bool fused = false;        // one pair evaluation feeds both sides (W per entry)
  int planned_mode = -1;     // tuning.adjoint_mode the plan was made for
  // captured passes (tuning.use_graph) and what they were captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  const void *graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int graph_cull = -1, graph_wpc = -1;
  int64_t nruns = 0;         // source leaves among the sorted particle segments
  int64_t *run_first = nullptr;
  int64_t *fw_first = nullptr, *tr_first = nullptr;
  int32_t *key = nullptr, *val = nullptr, *seg_t = nullptr;
};

struct Scratch {
  DBuf a, b, c, d, e, f, g, h, cub;
  DBuf cms;        // tree build: the four exact sums of each node of a level
  DBuf cmfail;     // tree build: small-node sums left to the ordered path
  DBuf cmx;        // tree build: big nodes' sums done by the integer path (bit q)
  DBuf x1, y1, z1, m1, p1, x0, y0, z0, m0, p0;
  DBuf blocks, pblocks, rights;
  DBuf cnode, pnode;  // node (BFS id) of each centre-of-mass / partition block
  DBuf pmask;         // partition flags (key < pivot), 64 per word, per block
  DBuf counters, qctr;
  DBuf host_in[10];
};

// Buffers of the sub-cycle steps (drift, resort) and the device-resident
// particle arrays hsr_subcycle keeps for host-pointer callers.
struct SubState {
  DBuf keys;       // cell keys and indices of the sort (16 B per particle)
  DBuf stage;      // gather staging of the ten arrays
  DBuf arr[10];    // x y z vx vy vz mass phi id mask on the device
  DBuf xcnt;       // refresh: per-tile category counts, their offsets, totals
  DBuf xstage;     // refresh: the exchanged particles before the resort
  DBuf xout;       // refresh: device output of host-pointer calls
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double drift_ms = 0, resort_ms = 0, subcycle_ms = 0, cic_ms = 0, inverse_cic_ms = 0;
  double refresh_ms = 0;
  int64_t refresh_alive = 0, refresh_dead = 0;
  float drift_factor = 0.f, kick_coeff = 0.f;  // last map1 / map2 factors
};

struct hsr_ctx {
  hsr_params p;
  ForceLawDev law;
  DBuf law_bins;  // LAWK_INTERP table on the device
  DBuf f64_at;    // FP64-at-positions source records and node masses
  DBuf f64_g;     // FP64-at-positions VJP with its node part
  hsr_tuning tuning;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8];
  int device = 0;
  int num_cus = 0;
  TreeState tree;
  ListState lists;
  AdjState adj;
  Scratch s;
  SubState sub;
  // caller arrays of the last build (for hsr_kick)
  float *user_v[3] = {nullptr, nullptr, nullptr};
  float *user_mass = nullptr;
  float *user_dv[3] = {nullptr, nullptr, nullptr};  // device velocity arrays
  int user_flags = 0;
  DBuf dev_v;  // device copy of velocities for host-pointer builds
  DBuf force;  // float3 SoA force [3][n]
  bool built = false, listed = false;
  double build_ms = 0, lists_ms = 0, force_ms = 0, kernel_ms = 0, adjoint_ms = 0;
  bool kernel_timed = false;
  // force phase captured as a hipGraph (tuning.use_graph)
  int64_t lists_gen = 0;
  int64_t evaluated_pairs = 0;  // last force phase, culled entries excluded
  int64_t unmasked_pairs = 0;   // of which unit-mass entries inside the cutoff
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int64_t graph_gen = -1;       // -1: the captured force phase is stale
  uintptr_t graph_sig[20] = {}; // what the captured force phase was built on
  int64_t graph_captures = 0, graph_launches = 0;
};

void hsr_set_error(const std::string &msg);

#define HSR_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      hsr_set_error(std::string("HIP error: ") + hipGetErrorString(e_) +  \
                    " at " + __FILE__ + ":" + std::to_string(__LINE__));  \
      return HSR_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

// launchers (defined in the .hip files)
hsr_status tree_build_device(hsr_ctx *c, int64_t nboxes,
                             const int64_t *box_off, const float *min_loc,
                             const float *max_loc, const float *x,
                             const float *y, const float *z, const float *m);
hsr_status lists_build_device(hsr_ctx *c, const float *min_force,
                              const float *max_force);
hsr_status force_plan(hsr_ctx *c);
// the leaf kernel's shape (targets per lane, entries per unrolled step,
// entries per LDS tile) for a tuning, 0 fields taking the defaults; whether
// that shape is built for a law; the built shapes as text
void force_shape(const hsr_tuning &t, int *r, int *u, int *tile);
bool force_shape_built(int law_kind, bool guard, int r, int u, int tile);
std::string force_shapes_built();
hsr_status force_run(hsr_ctx *c, float *ax, float *ay, float *az);
hsr_status kick_run(hsr_ctx *c, float fcoeff, float *vx, float *vy, float *vz,
                    const float *mass_sorted);
hsr_status adjoint_run(hsr_ctx *c, const float *lx, const float *ly,
                       const float *lz, float *gx, float *gy, float *gz);
hsr_status forces_f64_run(hsr_ctx *c, double *ax, double *ay, double *az, double *mom);
hsr_status forces_f64_at_run(hsr_ctx *c, const double *x, const double *y, const double *z,
                             double *ax, double *ay, double *az);
hsr_status adjoint_f64_at_run(hsr_ctx *c, const double *x, const double *y, const double *z,
                              const double *lx, const double *ly, const double *lz, double *gx,
                              double *gy, double *gz);
hsr_status leaf_momentum_run(hsr_ctx *c, const float *ax, const float *ay, const float *az,
                             double *mom);
hsr_status drift_run(hsr_ctx *c, int64_t n, float k, float *x, float *y, float *z,
                     const float *vx, const float *vy, const float *vz);
hsr_status resort_run(hsr_ctx *c, int64_t n, const int32_t ng[3], float *f[8],
                      int64_t *id, uint16_t *mask, int64_t *inside);
hsr_status fill_run(hsr_ctx *c, int64_t n, float v, float *a);
// Allocation-free resort over the library's static scratch array: the
// device's scratch address is looked up once per device at context
// creation (noalloc_scratch_init, device current).
hsr_status noalloc_scratch_init(int device);
hsr_status resort_noalloc_run(hsr_ctx *c, int64_t n, const int32_t ng[3], float *f[8],
                              int64_t *id, uint16_t *mask, int64_t *inside);

// Waits for the stream by polling it for a while before blocking: the
// tree build reads a few counters back once per level and the list build a
// few times, and a blocking wait wakes the host tens of microseconds after
// the copy has landed.
inline hipError_t stream_wait(hipStream_t s) {
  for (int i = 0; i < (1 << 20); ++i) {
    const hipError_t e = hipStreamQuery(s);
    if (e != hipErrorNotReady) return e;
  }
  return hipStreamSynchronize(s);
}

// The ten particle arrays (x y z vx vy vz mass phi, id, mask); entries may
// be null where an operation allows it.
struct PArrays {
  float *f[8];
  int64_t *id;
  uint16_t *mask;
};
// Stable cell sort of n particles from `in` into `out` (distinct arrays).
hsr_status resort_into(hsr_ctx *c, int64_t n, const int32_t ng[3], const PArrays &in,
                       const PArrays &out, int64_t *inside);

// One-rank overload refresh, in two halves so the caller can check the
// output size: refresh_count classifies the n particles (x, y, z of `in`)
// and returns the refreshed total and the alive count; refresh_finish packs
// them into staging and resorts them into `out` (room for n_out).
hsr_status refresh_count(hsr_ctx *c, const hsr_domain *dom, float anow, int64_t n,
                         const float *x, const float *y, const float *z, int64_t *n_out,
                         int64_t *alive);
hsr_status refresh_finish(hsr_ctx *c, const hsr_domain *dom, float anow, int64_t n,
                          const PArrays &in, int64_t n_out, const PArrays &out,
                          int64_t *inside);
void domain_derive(hsr_domain *d);
hsr_status cic_run(hsr_ctx *c, int64_t n, const float *x, const float *y, const float *z,
                   const int32_t ng[3], float gpscal, float *rho);
hsr_status inverse_cic_run(hsr_ctx *c, int64_t n, const float *x, const float *y,
                           const float *z, const int32_t ng[3], const float *grid, float tau,
                           float fscal, float *v);
