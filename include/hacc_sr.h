/*
 * hacc_sr.h -- C ABI of the MI355X short-range force library (libhaccsr.so).
 *
 * One context per host thread and device.  Every entry point returns an
 * hsr_status (0 = success); hsr_last_error() gives the message of the last
 * failure on the calling thread.  There is no CPU fallback: a context can
 * only be created on a HIP device.
 *
 * Reference surface replaced (BenWibking/hacc-coral, src/halo_finder):
 *   RCBForceTree<TDPTS>::RCBForceTree(...)   RCBForceTree.h:96-124,
 *                                             RCBForceTree.cxx:336-450
 *       = hsr_create + hsr_build + hsr_lists + hsr_kick (+ hsr_destroy),
 *         TDPTS 1 (RCBMonopoleForceTree) and 12 (RCBQuadrupoleForceTree)
 *   ForceLaw / ForceLawSR / FGridEvalPoly    ForceLaw.h:95-127
 *       = hsr_params.law (+ coefficients)
 *   extern "C" cm(...)                        RCBForceTree.cxx:514-516
 *       = device centre of mass inside hsr_build
 *   extern "C" Step16_int(...)                RCBForceTree.cxx:572
 *       = device leaf kernel inside hsr_forces / hsr_kick
 * and its caller in the sub-cycle (src/cpu):
 *   Particles::map1 / resortParticles         Particles.cxx:732-758, 402-488
 *       = hsr_drift / hsr_resort
 *   Particles::map2 (RCB monopole branch)     Particles.cxx:1205-1280
 *       = hsr_map2
 *   Particles::subCycle                       Particles.cxx:1176-1201
 *       = hsr_subcycle
 *   Particles::cic / inverse_cic              Particles.cxx:589-714
 *       = hsr_cic / hsr_inverse_cic
 *   MC3Extras refresh: Particles::copyAliveIntoVectors, ParticleExchange,
 *   copyAllFromVectors, resortParticles       MC3Extras.cxx:330-338,660-700,
 *                                             ParticleExchange.cxx:488-728
 *       = hsr_domain_init + hsr_refresh (one-rank decomposition)
 */
#ifndef HACC_SR_H
#define HACC_SR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int hsr_status;
enum {
  HSR_OK = 0,
  HSR_ERR_ARG = 1,        /* invalid argument */
  HSR_ERR_HIP = 2,        /* HIP runtime error */
  HSR_ERR_UNSUPPORTED = 3,/* configuration the device path does not implement */
  HSR_ERR_STATE = 4,      /* call out of order (e.g. kick before build) */
  HSR_ERR_NOMEM = 5,      /* device allocation failed */
  HSR_ERR_NUMERIC = 6     /* an input the exact build cannot represent */
};

/* force laws (ForceLaw.h:95-127, BGQStep16.c:167-187) */
enum {
  HSR_LAW_STEP16 = 0,    /* (r^2+rsm^2)^-3/2 - P5(r^2), r^2 < fsrrmax^2, r^2 > 0 */
  HSR_LAW_SR_POLY5 = 1,  /* ForceLawSR(FGridEvalPoly), POLY_ORDER 5 */
  HSR_LAW_SR_POLY6 = 2,  /* ForceLawSR(FGridEvalPoly), POLY_ORDER 6 (stock) */
  HSR_LAW_NEWTON = 3,    /* ForceLawNewton, 1/r^3 (not on the device) */
  HSR_LAW_SR_FIT = 4,    /* ForceLawSR(FGridEvalFit): the fitted grid force
                            FGrid::fgor (ForceLaw.cxx:39-51) per pair; the
                            reference's default law (Particles.cxx:173-174) */
  HSR_LAW_SR_INTERP = 5  /* ForceLawSR(FGridEvalInterp): FGrid::fgor tabulated
                            on n_interp points of r^2 and interpolated linearly
                            (ForceLaw.cxx:54-67,145-171); the table comes from
                            the caller's FGridEvalInterp (hsr_set_law_table) */
};

/* pointer-location flags */
enum {
  HSR_HOST_PTRS = 0,     /* particle arrays are host memory */
  HSR_DEVICE_PTRS = 1    /* particle arrays are device memory on the context's device */
};

typedef struct {
  int device;            /* HIP device ordinal */
  int tdpts;             /* pseudo-particles per node: 1 (RCBMonopoleForceTree)
                            or 12 (RCBQuadrupoleForceTree; no hsr_adjoint) */
  int law;               /* HSR_LAW_* */
  int n_interp;          /* HSR_LAW_SR_INTERP: table size nInterp (>= 2) */
  int64_t nd;            /* leaf size ("nDirect", ctor arg nd) */
  int64_t ds;            /* depth safety (ctor arg ds): initial node reservation */
  int64_t tmin;          /* ctor arg tmin (task threshold; no effect on results) */
  float fsrrmax;         /* ctor arg fsm */
  float rsm;             /* ctor arg r (Plummer softening) */
  float oa;              /* ctor arg oa (opening angle) */
  float tan_oa;          /* tanf(oa) as the host libm evaluates it (RCBForceTree.cxx:381) */
  float ppc;             /* ctor arg ppc (pseudo-particle contraction, 0.9) */
  float reserved1[3];
} hsr_params;

/* Tree node in the reference's layout and canonical numbering
 * (RCBForceTree.h:131-147 with TDPTS = 1; ids as the unthreaded build
 * allocates them, RCBForceTree.cxx:803-825). */
typedef struct {
  int64_t count, offset, cl, cr;
  float ppm[1];
  float tdr;
  float xmin[3], xmax[3], xc[3];
} hsr_node;

typedef struct hsr_ctx hsr_ctx;

const char *hsr_last_error(void);
const char *hsr_version(void);

hsr_status hsr_create(hsr_ctx **ctx, const hsr_params *p);
hsr_status hsr_destroy(hsr_ctx *ctx);

/* The table of an HSR_LAW_SR_INTERP context: FGridEvalInterp's r2() and f()
 * arrays of nInterp entries (ForceLaw.h:73-91, filled by
 * FGrid::fgor_r2_interp, ForceLaw.cxx:54-67), copied to the device; n must
 * equal params.n_interp and r2 must ascend from r2[0].  Required before the
 * first hsr_forces / hsr_kick of such a context (HSR_ERR_STATE otherwise);
 * other laws refuse it. */
hsr_status hsr_set_law_table(hsr_ctx *ctx, const float *r2, const float *f, int64_t n);

/* The HIP stream the context launches on (hipStream_t as void*). */
hsr_status hsr_get_stream(hsr_ctx *ctx, void **stream);

/* Tree build (RCBForceTree.cxx:893-918 with cm, partition, pseudo-particles).
 * Permutes all ten arrays in place exactly as the reference does and
 * records the tree.  min_loc/max_loc: the tree box (root bbox seed). */
hsr_status hsr_build(hsr_ctx *ctx, int64_t count, float *x, float *y,
                     float *z, float *vx, float *vy, float *vz, float *mass,
                     float *phi, int64_t *id, uint16_t *mask,
                     const float min_loc[3], const float max_loc[3],
                     int ptr_flags);

/* Interaction lists of every leaf inside the force box
 * (RCBForceTree.cxx:924-1197). */
hsr_status hsr_lists(hsr_ctx *ctx, const float min_force[3],
                     const float max_force[3]);

/* Raw short-range force F_i = sum_j m_j f(r_ij^2) (x_j - x_i) for particles
 * of the listed leaves, in tree (sorted) order; zero for other particles. */
hsr_status hsr_forces(hsr_ctx *ctx, float *ax, float *ay, float *az,
                      int ptr_flags);

/* The kick of the reference constructor: v += fcoeff * m_i * F_i on the
 * particles of the listed leaves, applied to the velocity arrays given to
 * hsr_build (nbody1, RCBForceTree.cxx:599-619). */
hsr_status hsr_kick(hsr_ctx *ctx, float fcoeff);

/* Vector-Jacobian product of the force with respect to positions:
 * g = d(sum_i lambda_i . F_i)/dx for the current tree and lists, including
 * the pseudo-particle centre-of-mass dependence (F as returned by
 * hsr_forces).  lambda and g are in tree order; g covers every particle
 * (zero where no listed pair touches it).  No counterpart in the reference:
 * the derivative of its pair law (BGQStep16.c:167-187) with the tree
 * topology and lists held fixed. */
hsr_status hsr_adjoint(hsr_ctx *ctx, const float *lx, const float *ly,
                       const float *lz, float *gx, float *gy, float *gz,
                       int ptr_flags);

/* ---------------------------------------------------------- introspection */
typedef struct {
  int64_t count;          /* particles */
  int64_t nodes;          /* tree.size() */
  int64_t leaves;         /* leaves with count > 0 */
  int64_t empty_leaves;   /* orphan/empty leaves (printStats) */
  int64_t max_ppn;        /* largest leaf */
  int64_t listed_leaves;  /* leaves inside the force box */
  int64_t segments;       /* list segments over listed leaves */
  int64_t entries;        /* list entries over listed leaves */
  int64_t pairs;          /* sum over listed leaves of count * list length */
  int64_t max_depth;
  double build_ms, lists_ms, force_ms; /* device time of the last calls */
  double kernel_ms;       /* the leaf-kernel launch alone (last force/kick) */
  double adjoint_ms;      /* device time of the last hsr_adjoint */
  int64_t evaluated_pairs;/* pairs the last force/kick evaluated: list pairs
                             minus those of entries culled beyond the cutoff */
  double drift_ms;        /* device time of the last hsr_drift (hsr_subcycle:
                             all its drifts) */
  double resort_ms;       /* the same for hsr_resort */
  double subcycle_ms;     /* host wall time of the last hsr_subcycle */
  double drift_factor;    /* prefactor * tau of the last drift (a float) */
  double kick_coeff;      /* c of the last map2 kick (a float) */
  double cic_ms;          /* device time of the last hsr_cic */
  double inverse_cic_ms;  /* device time of the last hsr_inverse_cic */
  double refresh_ms;      /* device time of the last hsr_refresh */
  int64_t refresh_alive;  /* particles it kept alive */
  int64_t refresh_dead;   /* overload copies it appended */
  int64_t unmasked_pairs; /* of evaluated_pairs, those of unit-mass entries
                             whose every pair lies inside the cutoff (the leaf
                             kernel skips their mask and mass) */
  int64_t graph_captures; /* force-phase graphs captured (tuning.use_graph):
                             a new one only when a buffer address, the
                             particle count or the settings change */
  int64_t graph_replays;  /* force phases run as a replay of such a graph */
} hsr_stats;

hsr_status hsr_get_stats(hsr_ctx *ctx, hsr_stats *st);
/* Tree in canonical numbering (host buffer of stats.nodes entries). */
hsr_status hsr_get_nodes(hsr_ctx *ctx, hsr_node *out, int64_t cap);
/* Pseudo-particle masses, TDPTS per node in canonical node order
 * (RCBForceTree.h:139, ppm[TDPTS]); cap counts floats.  hsr_node.ppm[0]
 * holds the first of them. */
hsr_status hsr_get_pseudo_masses(hsr_ctx *ctx, float *ppm, int64_t cap);
/* perm[k] = input index of the particle now at position k. */
hsr_status hsr_get_perm(hsr_ctx *ctx, int64_t *perm, int64_t cap);
/* Lists in the reference's order: listed leaves in the order
 * calcInternodeForces visits them; per leaf, segments (a, b): particle
 * range (offset a, count b >= 0) or pseudo-particle of node a (b = -1). */
hsr_status hsr_get_lists(hsr_ctx *ctx, int64_t *leaves, int64_t *seg_ptr,
                         int64_t *seg_a, int64_t *seg_b, int64_t *entries,
                         int64_t cap_leaves, int64_t cap_segs);

/* ------------------------------------------------------------- ensemble */
/* Independent boxes in one context: box b owns particles
 * [offsets[b], offsets[b+1]) of the arrays; each is built as its own tree,
 * and all boxes' lists are processed by one persistent force launch.
 * Boxes' tree/force boxes are given per box (3 floats each). */
hsr_status hsr_build_many(hsr_ctx *ctx, int64_t nboxes, const int64_t *offsets,
                          float *x, float *y, float *z, float *vx, float *vy,
                          float *vz, float *mass, float *phi, int64_t *id,
                          uint16_t *mask, const float *min_loc,
                          const float *max_loc, int ptr_flags);
hsr_status hsr_lists_many(hsr_ctx *ctx, const float *min_force,
                          const float *max_force);

/* ------------------------------------------------------ FP64 checks */
/* No counterpart in the reference: FP64 evaluations for the precision and
 * stability studies, on the device, over the current tree and lists.
 *
 * hsr_forces_f64: a_i = sum_j m_j f(r2) d in double (float positions and
 * masses promoted; pseudo-particles at their float centre with mass ppm;
 * f = (r2 + rsm^2)^-3/2 - P(r2) with the law's float coefficients promoted,
 * r2 > 0 and r2 < cutoff as in BGQStep16.c:167-187).  Tree order, like
 * hsr_forces; particles outside the listed leaves get 0.  If leaf_mom is not
 * NULL it receives sum_i m_i a_i of every listed leaf (3 doubles per leaf, in
 * hsr_get_lists leaf order).
 *
 * hsr_leaf_momentum: sum_i m_i a_i per listed leaf, accumulated in double, of
 * float accelerations a in tree order (as hsr_forces returns them). */
hsr_status hsr_forces_f64(hsr_ctx *ctx, double *ax, double *ay, double *az,
                          double *leaf_mom, int ptr_flags);
hsr_status hsr_leaf_momentum(hsr_ctx *ctx, const float *ax, const float *ay,
                             const float *az, double *leaf_mom, int ptr_flags);

/* FP64 oracle of the differentiation checks (no counterpart in the
 * reference): the force and its vector-Jacobian product at given double
 * positions with the tree and lists of the last hsr_build / hsr_lists
 * frozen.  Positions are in tree order; pseudo-particles sit at the FP64
 * centre of mass of their members' given positions with mass ppm; the law is
 * the context's polynomial in r2 in double (the Step16 law: r2 < fsrrmax^2,
 * pairs at r2 = 0 skipped only when rsm = 0).  hsr_adjoint_f64_at gives
 * g = d(sum_i lambda_i . F_i)/dx including the pseudo-particle centres.
 * Device pointers only, arrays of stats.count doubles; a single monopole
 * tree (hsr_build, TDPTS 1). */
hsr_status hsr_forces_f64_at(hsr_ctx *ctx, const double *x, const double *y,
                             const double *z, double *ax, double *ay, double *az);
hsr_status hsr_adjoint_f64_at(hsr_ctx *ctx, const double *x, const double *y,
                              const double *z, const double *lx, const double *ly,
                              const double *lz, double *gx, double *gy, double *gz);

/* ------------------------------------------------------------ sub-cycle */
/* The drift of the sub-cycle, Particles::map1(pp, tau, adot)
 * (Particles.cxx:732-758): x += prefactor * tau * v over `count` particles
 * with prefactor = 1 / (alpha * adot * powf(pp, 1 + 1/alpha)) evaluated on
 * the host as there (alpha = m_alpha) and every element rounded as the
 * reference's float expression x + (prefactor * tau) * v. */
hsr_status hsr_drift(hsr_ctx *ctx, int64_t count, float *x, float *y, float *z,
                     const float *vx, const float *vy, const float *vz,
                     float pp, float tau, float adot, float alpha,
                     int ptr_flags);

/* Particles::resortParticles (Particles.cxx:402-488): stable counting sort
 * of the particles by unit cell, cell = (floor x * ng[1] + floor y) * ng[2] +
 * floor z for particles inside [0, ng) per dimension, all others after the
 * last cell, input order kept within a cell.  Permutes the ten arrays in
 * place (vx .. mask may be NULL) and writes the number of particles inside
 * the grid (m_Np_last) to *inside.  ng[0]*ng[1]*ng[2] must stay below 2^31. */
hsr_status hsr_resort(hsr_ctx *ctx, int64_t count, const int32_t ng[3],
                      float *x, float *y, float *z, float *vx, float *vy,
                      float *vz, float *mass, float *phi, int64_t *id,
                      uint16_t *mask, int64_t *inside, int ptr_flags);

/* hsr_resort without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * hsr_resort with HSR_DEVICE_PTRS, bit for bit, but the cell keys, the radix
 * sort's storage and the staging copy live in one static device array of
 * HSR_RESORT_NOALLOC_SCRATCH_BYTES that the library holds on every device it
 * runs on until the process ends; the call makes no hipMalloc or hipFree.
 * Limits: device pointers only (ptr_flags must be HSR_DEVICE_PTRS) and at
 * most HSR_RESORT_NOALLOC_MAX particles (HSR_ERR_ARG beyond, before any
 * launch).  Every call on a device shares that device's array, so calls that
 * overlap on one device (other contexts, other streams or threads) race: the
 * caller serialises them. */
#define HSR_RESORT_NOALLOC_MAX ((int64_t)1 << 22)
#define HSR_RESORT_NOALLOC_SCRATCH_BYTES \
  (58ull * (1ull << 22) + 512ull + (64ull << 20))
hsr_status hsr_resort_noalloc(hsr_ctx *ctx, int64_t count, const int32_t ng[3],
                              float *x, float *y, float *z, float *vx, float *vy,
                              float *vz, float *mass, float *phi, int64_t *id,
                              uint16_t *mask, int64_t *inside, int ptr_flags);

/* Particles::subCycle with its short-range kick Particles::map2
 * (Particles.cxx:1176-1201, 1205-1280; RCB monopole tree): for each of nsub
 * sub-steps, drift by stepFraction * tau2 (map1), then resort, set every
 * mass to 1, build the tree over the particles inside the grid with the tree
 * box [0, max ng) and list and kick the leaves inside the force box
 * [edge, ng - edge) with c = gpscal^3/(4 pi) * fscal * tau * stepFraction,
 * then drift again.  The particles stay on the device for the whole cycle
 * (host arrays are copied in once and out once).  Tree and kernel parameters
 * (nd, fsrrmax, rsm, oa, law) are the context's. */
typedef struct {
  int nsub;              /* m_nsub (stepFraction = 1.0 / nsub) */
  int skip_stream;       /* m_skipStreamQ: no drifts */
  int skip_kick;         /* m_skipKickSRQ: no map2 */
  int32_t ng[3];         /* Domain::ng_local_total */
  float alpha;           /* m_alpha */
  float gpscal;          /* m_gpscal */
  float edge;            /* m_edge */
  float reserved0;
  double pp, tau, tau2, adot, fscal;  /* the TimeStepper's values */
} hsr_subcycle_params;

/* Particles::map2 alone (RCB monopole branch, Particles.cxx:1205-1280):
 * resort, unit masses, tree, lists and kick as one sub-step of hsr_subcycle
 * without its drifts, with the caller's stepFraction (sp->nsub, skip_* and
 * the drift's TimeStepper values are not read). */
hsr_status hsr_map2(hsr_ctx *ctx, const hsr_subcycle_params *sp,
                    double stepFraction, int64_t count, float *x, float *y,
                    float *z, float *vx, float *vy, float *vz, float *mass,
                    float *phi, int64_t *id, uint16_t *mask, int64_t *inside,
                    int ptr_flags);

hsr_status hsr_subcycle(hsr_ctx *ctx, const hsr_subcycle_params *sp,
                        int64_t count, float *x, float *y, float *z,
                        float *vx, float *vy, float *vz, float *mass,
                        float *phi, int64_t *id, uint16_t *mask,
                        int64_t *inside, int ptr_flags);

/* ------------------------------------------------------------ PM steps */
/* Particles::cic (Particles.cxx:589-641): cloud-in-cell deposit of
 * c = gpscal^3 per particle into rho, which holds ng[0]*ng[1]*ng[2] + 1
 * floats (row-major cells, then the slot that collects corners outside the
 * grid); rho is zeroed first.  Bit-exact with the reference's sequential
 * order of float additions into each cell. */
hsr_status hsr_cic(hsr_ctx *ctx, int64_t count, const float *x, const float *y,
                   const float *z, const int32_t ng[3], float gpscal, float *rho,
                   int ptr_flags);

/* Particles::inverse_cic(tau, fscal, comp) (Particles.cxx:645-714) for one
 * component: v += fscal * tau * (CIC interpolation of grid at each
 * particle), grid holding ng[0]*ng[1]*ng[2] + 1 floats whose last slot is
 * read as 0 (the reference zeroes it; here it is left untouched). */
hsr_status hsr_inverse_cic(hsr_ctx *ctx, int64_t count, const float *x,
                           const float *y, const float *z, const int32_t ng[3],
                           const float *grid, float tau, float fscal, float *v,
                           int ptr_flags);

/* ------------------------------------------------------ overload refresh */
/* The rank's domain (Domain::initialize, Domain.cxx:34-84) for a one-rank
 * decomposition.  The caller sets rL, oL_in and ng (the input deck's rL, oL
 * and ng); hsr_domain_init derives the rest with the reference's float and
 * double operations: phys2grid_pos = ng/rL, grid2phys_pos = rL/ng,
 * phys2grid_vel = ng/(100 rL), grid2phys_vel = 100 rL/ng, ng_overload =
 * ceilf(oL_in * phys2grid_pos), oL = ng_overload * grid2phys_pos and
 * ng_local_total = ng + 2 ng_overload + 1 (the grid of hsr_resort, hsr_map2
 * and hsr_subcycle). */
typedef struct {
  float rL;              /* box side, physical units */
  float oL_in;           /* requested overload width, physical units */
  int32_t ng;            /* grid points per side */
  int32_t ng_overload;   /* derived */
  int32_t ng_local_total;/* derived */
  float oL;              /* derived: overload width actually used */
  float phys2grid_pos, grid2phys_pos, phys2grid_vel, grid2phys_vel; /* derived */
} hsr_domain;

hsr_status hsr_domain_init(hsr_domain *dom);

/* The particle refresh of MC3Extras::particleExtras (MC3Extras.cxx:330-338,
 * 660-700) on one rank, with the particles on the device throughout:
 *  1. Particles::copyAliveIntoVectors (Particles.cxx:945-1030): local grid
 *     units to physical ((x + corner) * grid2phys_pos, v * grid2phys_vel /
 *     anow / anow, corner = -ng_overload) and keep, in order, the particles
 *     with every coordinate in [0, rL);
 *  2. ParticleExchange::exchangeParticles (ParticleExchange.cxx:488-728)
 *     with every neighbour this rank: a kept particle outside the open
 *     interior (oL, rL - oL)^3 is copied, for each of the 26 neighbours in
 *     the order of BasicDefinition.h:224-264 whose closed region holds it,
 *     with its position shifted by -rL, 0 or +rL per dimension; copies are
 *     appended neighbour by neighbour, in particle order within one;
 *  3. Particles::copyAllFromVectors (Particles.cxx:298-366): every mass set
 *     to 1, back to local grid units (x * phys2grid_pos - corner,
 *     v * phys2grid_vel * anow * anow);
 *  4. Particles::resortParticles on the ng_local_total grid (as hsr_resort).
 * The ten arrays hold *count particles on entry and have room for capacity;
 * on return *count is the refreshed number and *inside the particles inside
 * the grid.  When the refreshed number exceeds capacity the call fails with
 * HSR_ERR_ARG, leaves the arrays as they were and writes the number it needs
 * to *count.  With more ranks step 2 sends the 26
 * packed sets to the neighbour ranks; that transport is not part of this
 * library. */
hsr_status hsr_refresh(hsr_ctx *ctx, const hsr_domain *dom, float anow,
                       int64_t *count, int64_t capacity, float *x, float *y,
                       float *z, float *vx, float *vy, float *vz, float *mass,
                       float *phi, int64_t *id, uint16_t *mask,
                       int64_t *inside, int ptr_flags);

/* ------------------------------------------------------------- tuning */
typedef struct {
  int i_per_lane;     /* leaf kernel shape: i-particles held per lane (0 = 2);
                         with unroll and tile one of the built shapes
                         (i_per_lane, unroll, tile): (2,2,256) the default,
                         (2,1,256) (2,4,256) (1,2,256) (1,4,256) (4,1,256)
                         (2,2,128) (2,2,512), the non-default ones for the
                         polynomial laws with rsm > 0; sums in another order:
                         rounding only */
  int tile;           /* j entries staged per LDS tile (0 = 256) */
  int chunk;          /* maximum list entries per work item: 0 (= 8192) or
                         tile..2^30; a change re-plans lists already built
                         (on failure the old tuning and plan are restored
                         where possible; otherwise the lists must be built
                         again and calls that need them give HSR_ERR_STATE) */
  int waves_per_cu;   /* persistent waves per CU, 0..256 (0 = 32) */
  int use_graph;      /* capture the force phase in a hipGraph and replay it */
  int adjoint_mode;   /* hsr_adjoint passes: 0 automatic, 1 fused (one pair
                         evaluation feeds target and source; 12 B of device
                         memory per list entry), 2 separate target and source
                         passes (automatic picks fused when that memory is
                         free) */
  int cull_off;       /* 1: evaluate every list entry; 0 (default): skip entries
                         farther than the cutoff from the box of the targets
                         of a pass, whose pairs are all masked */
  int tree_order_targets; /* 1: take a leaf's targets in tree order; 0
                         (default): sorted along the leaf's longest edge, so
                         that each pass of 128 targets spans a slab and culls
                         more entries (sums in another order: rounding only) */
  int force_pass;     /* targets per pass of the leaf kernel: 128 (0 = default)
                         or 64 (smaller target boxes cull more list entries;
                         sums in another order: rounding only); 64 and 32
                         times i_per_lane in general */
  int64_t max_plan_bytes; /* cap on the device memory of the force plan (work
                         items, their sort keys, chunk partials); 0: no cap.
                         A plan above it fails with HSR_ERR_NOMEM before any
                         launch (so does one of 2^31 work items or more) */
  int unroll;         /* leaf kernel shape: j entries per unrolled step (0 = 2) */
  int chain_hops;     /* tree build: stretch steps a partition swap chain takes
                         before the pointer-jumping pass resolves it, 0..4096
                         (0 = 256); the tree is the same for every value */
} hsr_tuning;

/* Out-of-range fields give HSR_ERR_ARG and leave the tuning unchanged. */

hsr_status hsr_set_tuning(hsr_ctx *ctx, const hsr_tuning *t);
hsr_status hsr_get_tuning(hsr_ctx *ctx, hsr_tuning *t);

#ifdef __cplusplus
}
#endif

#endif /* HACC_SR_H */
