/*
 * hsr_oracle.h -- CPU restatement of HACC's short-range RCB force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, the smoke entry point and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Every function restates the reference algorithm (BenWibking/hacc-coral,
 * src/halo_finder) in plain C with the same float/double operation order:
 *   hsro_cm              BGQCM.c:181-212 (generic branch)
 *   hsro_partition       RCBForceTree.cxx:623-672
 *   hsro_build           RCBForceTree.cxx:690-918 (unthreaded build order)
 *   hsro_leaf_order      RCBForceTree.cxx:1151-1197
 *   hsro_list            RCBForceTree.cxx:924-1140
 *   hsro_force_*         RCBForceTree.cxx:575-620 / BGQStep16.c:167-187
 *   hsro_pppts, TDPTS 12 RCBForceTree.cxx:229-272, 519-569, 1051-1060
 *   hsro_cell_sort_perm  Particles.cxx:402-488 (resortParticles)
 *   hsro_map1_factor, hsro_drift   Particles.cxx:732-758 (map1)
 *   hsro_map2_coeff      Particles.cxx:1230-1233 (map2)
 *   hsro_cic, hsro_inverse_cic     Particles.cxx:589-714
 *   hsro_domain_init     Domain.cxx:34-84 (one rank)
 *   hsro_refresh         MC3Extras.cxx:330-338, 660-700 with
 *                        Particles.cxx:298-366, 762-830, 945-1030 and
 *                        ParticleExchange.cxx:103-728 (src/simulation,
 *                        src/cpu, src/halo_finder)
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   cm            bit-exact against the reference's own BGQCM.c compiled
 *                 into oracle/_ref, on every node range of built trees.
 *   force laws    bit-exact against the reference's ForceLaw.cxx (_ref).
 *   partition     stable-filter + swap-chain invariants, brute-force replay.
 *   tree          tiling/tight-bbox/split-rule/leaf-size invariants.
 *   lists         mass-accounting invariant and oa->0 direct-sum limit.
 *   kernels       FP64 brute force over the same lists.
 *   TDPTS = 12    the design is the icosahedron's 12 vertices; every node's
 *                 pseudo-particles reproduce its members' mass, dipole and
 *                 traceless quadrupole about xc; the tree force is closer
 *                 to the direct sum than the monopole tree's.
 *   cell sort     numpy's stable argsort of the cell keys.
 *   CIC           weights sum to one per particle (mass conservation),
 *                 linear fields interpolate exactly, a particle on a grid
 *                 point deposits into that cell alone.
 *   map1, map2    closed forms (prefactor 1 exactly, double-precision
 *                 formula to float rounding, two half drifts vs one).
 *   domain        known answers from exact rational arithmetic.
 *   refresh       the copies are exactly the periodic images k*rL,
 *                 k in {-1,0,1}^3 \ 0, of the alive particles whose
 *                 shift is allowed by the closed overload shell, enumerated
 *                 by brute force; neighbour order from the enum's names.
 */
#ifndef HSR_ORACLE_H
#define HSR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSRO_MAX_TDPTS 1

/* Node layout follows RCBForceTree<TDPTS>::TreeNode (RCBForceTree.h:131-147)
 * with TDPTS = 1 (RCBMonopoleForceTree); with TDPTS = 12
 * (RCBQuadrupoleForceTree) the twelve masses live in hsro_params.ppm_t. */
typedef struct {
  int64_t count, offset, cl, cr;
  float ppm[HSRO_MAX_TDPTS];
  float tdr;
  float xmin[3], xmax[3], xc[3];
} hsro_node;

typedef struct {
  int64_t nd;          /* leaf size: nodes with count <= nd are leaves */
  float fsrrmax;       /* short-range cutoff */
  float rsm;           /* Plummer softening */
  float tan_oa;        /* tanf(opening angle), evaluated on the host */
  float ppc;           /* pseudo-particle contraction factor (0.9) */
  float min_loc[3], max_loc[3];     /* tree box (root bbox seed) */
  float min_force[3], max_force[3]; /* force box */
  int32_t tdpts;       /* pseudo-particles per node: 1 (also when 0) or 12 */
  int32_t n_interp;    /* HSRO_LAW_SR_INTERP: table size nInterp (>= 2) */
  float *ppm_t;        /* tdpts > 1: pseudo masses, tdpts per node in node
                          order (the build writes them, lists and kernels
                          read them); node.ppm[0] then holds the first */
} hsro_params;

/* force-law kinds */
enum {
  HSRO_LAW_STEP16 = 0,   /* BGQStep16.c:167-187: P5 law, r2>0 guard, pow() in double */
  HSRO_LAW_SR_POLY5 = 1, /* ForceLawSR + FGridEvalPoly, POLY_ORDER 5 (ForceLaw.cxx:107-116,137-141,187-192) */
  HSRO_LAW_SR_POLY6 = 2, /* ForceLawSR + FGridEvalPoly, POLY_ORDER 6 (stock build, ForceLaw.cxx:96-105) */
  HSRO_LAW_NEWTON = 3,   /* ForceLawNewton (ForceLaw.h:105-110) */
  HSRO_LAW_SR_FIT = 4,   /* ForceLawSR + FGridEvalFit, the reference's default law
                            (Particles.cxx:173-174; ForceLaw.cxx:39-51,71-81) */
  HSRO_LAW_SR_INTERP = 5 /* ForceLawSR + FGridEvalInterp(nInterp)
                            (ForceLaw.cxx:54-67,145-171) */
};

/* BGQCM.c:181-212, generic branch. */
void hsro_cm(int64_t count, const float *xx, const float *yy, const float *zz,
             const float *mass, float *xmin, float *xmax, float *xc);

/* RCBForceTree.cxx:623-672; returns is (number of keys < pv). The first
 * three arrays are the key-rotated coordinate arrays. */
int64_t hsro_partition(int64_t n, float *xx, float *yy, float *zz, float *vx,
                       float *vy, float *vz, float *mass, float *phi,
                       int64_t *id, uint16_t *mask, float pv);

/* Full unthreaded build (RCBForceTree.cxx:893-918).  Permutes all ten arrays
 * in place.  nodes must hold cap entries; returns the node count or -1 when
 * cap is too small. */
int64_t hsro_build(const hsro_params *p, int64_t n, float *x, float *y,
                   float *z, float *vx, float *vy, float *vz, float *mass,
                   float *phi, int64_t *id, uint16_t *mask, hsro_node *nodes,
                   int64_t cap);

/* Leaves inside the force box in the order calcInternodeForces visits them
 * (RCBForceTree.cxx:1151-1197).  Writes node ids and the ancestor chain
 * needed by hsro_list into out; returns the count (or -1 if cap too small). */
int64_t hsro_leaf_order(const hsro_params *p, const hsro_node *nodes,
                        int64_t nnodes, int64_t *out, int64_t cap);

/* Interaction list of target leaf tl (RCBForceTree.cxx:924-1140), as
 * segments: seg_a[k] = particle offset and seg_b[k] = count for a particle
 * range, or seg_a[k] = node id and seg_b[k] = -1 for a pseudo-particle.
 * The target's own particles are the last segment.  Returns the segment
 * count (or -1 when cap is too small); *n_entries receives the list length. */
int64_t hsro_list(const hsro_params *p, const hsro_node *nodes,
                  int64_t nnodes, int64_t tl, int64_t *seg_a, int64_t *seg_b,
                  int64_t cap, int64_t *n_entries);

/* All lists for the leaves in `leaves` (CSR over segments). seg_ptr has
 * nleaves+1 entries.  Returns total segments or -1. Uses OpenMP. */
int hsro_set_threads(int n);
int64_t hsro_lists_all(const hsro_params *p, const hsro_node *nodes,
                       int64_t nnodes, const int64_t *leaves, int64_t nleaves,
                       int64_t *seg_ptr, int64_t *seg_a, int64_t *seg_b,
                       int64_t cap, int64_t *entries_per_leaf);

/* Float force of the generic kernel nbody1 (RCBForceTree.cxx:599-619) with
 * fcoeff = 1 and v = 0, i.e. the per-pair accumulation of the reference's
 * float CPU path, for every listed leaf.  ax/ay/az are indexed by sorted
 * particle position; entries of non-target particles are untouched. */
void hsro_force_f32(int law, const hsro_params *p, const hsro_node *nodes,
                    const float *x, const float *y, const float *z,
                    const float *m, const int64_t *leaves, int64_t nleaves,
                    const int64_t *seg_ptr, const int64_t *seg_a,
                    const int64_t *seg_b, float *ax, float *ay, float *az);

/* Step16 structure (BGQStep16.c:167-187): per-i accumulator a, then
 * v += fcoeff * a.  Same lists. */
void hsro_force_step16(const hsro_params *p, const hsro_node *nodes,
                       const float *x, const float *y, const float *z,
                       const float *m, const int64_t *leaves, int64_t nleaves,
                       const int64_t *seg_ptr, const int64_t *seg_a,
                       const int64_t *seg_b, float *ax, float *ay, float *az);

/* FP64 evaluation of the Step16 5th-order law over the same lists (inputs are the
 * float positions promoted to double).  Also returns, per listed leaf, the
 * FP64 momentum sum sum_i m_i F_i (3 doubles per leaf) if leaf_mom != 0. */
void hsro_force_f64(const hsro_params *p, const hsro_node *nodes,
                    const float *x, const float *y, const float *z,
                    const float *m, const int64_t *leaves, int64_t nleaves,
                    const int64_t *seg_ptr, const int64_t *seg_a,
                    const int64_t *seg_b, double *ax, double *ay, double *az,
                    double *leaf_mom);

/* Scalar law f(r2) as used by the kernels above, for pinning against the
 * reference's ForceLaw.cxx (returns f_over_r without the mass factors and,
 * for STEP16, with the r2 < fsrrmax^2 mask and r2 > 0 guard applied). */
float hsro_law_f32(int law, float r2, float rsm, float fsrrmax);
/* hsro_law_f32 over n values (n_interp: table size of HSRO_LAW_SR_INTERP). */
void hsro_law_f32_vec(int law, int64_t n, const float *r2, float *out, float rsm,
                      float fsrrmax, int n_interp);
/* Polynomial coefficients a0.. of a polynomial law (STEP16 and SR_POLY5: the
 * five-term set of BGQStep16.c:40-45,167 / ForceLaw.cxx:109-114; SR_POLY6:
 * ForceLaw.cxx:98-104) into out[7]; returns how many (0: not a polynomial law). */
int hsro_law_coeffs(int law, float *out);
/* FGrid::fgor_r2_interp (ForceLaw.cxx:54-67): r2[i], f[i] for i < n. */
void hsro_fgrid_table(int n, float *r2, float *f);

/* Brute-force FP64 short-range force on particle i from all particles j with
 * 0 < r2 < fsrrmax^2 (the oa -> 0 limit of the tree). */
void hsro_direct_f64(int64_t n, const float *x, const float *y, const float *z,
                     const float *m, const int64_t *targets, int64_t ntargets,
                     float rsm, float fsrrmax, double *ax, double *ay,
                     double *az);

/* Adjoint of the Step16 5th-order law over the same lists, in FP64 (there is no
 * reference adjoint; the mathematics is pinned by finite differences of
 * hsro_objective_f64).  J = sum_i lam_i . F_i with F as hsro_force_f64
 * defines it and pseudo-particles at the FP64 centre of mass of their
 * node's members; returns dJ/dx (n x 3, by sorted position). */
void hsro_adjoint_f64(const hsro_params *p, const hsro_node *nodes,
                      const double *x, const double *y, const double *z,
                      const double *m, const int64_t *leaves, int64_t nleaves,
                      const int64_t *seg_ptr, const int64_t *seg_a,
                      const int64_t *seg_b, const double *lx, const double *ly,
                      const double *lz, double *gx, double *gy, double *gz);

/* FP64 force at double positions (sorted order) with the tree topology and
 * lists frozen; pseudo-particles at the FP64 centre of mass of their
 * members, mass ppm.  Writes a (n x 3) for listed particles. */
void hsro_force_f64d(const hsro_params *p, const hsro_node *nodes,
                     const double *x, const double *y, const double *z,
                     const double *m, const int64_t *leaves, int64_t nleaves,
                     const int64_t *seg_ptr, const int64_t *seg_a,
                     const int64_t *seg_b, double *ax, double *ay, double *az);

double hsro_objective_f64(const hsro_params *p, const hsro_node *nodes,
                          const double *x, const double *y, const double *z,
                          const double *m, const int64_t *leaves,
                          int64_t nleaves, const int64_t *seg_ptr,
                          const int64_t *seg_a, const int64_t *seg_b,
                          const double *lx, const double *ly,
                          const double *lz);

/* Stable counting sort by unit cell (Particles.cxx:402-488): returns the
 * permutation perm[i] = destination of particle i and the number of
 * particles inside the ng^3 grid (the rest go to the tail). */
int64_t hsro_cell_sort_perm(int64_t n, const float *x, const float *y,
                            const float *z, const int32_t ng[3],
                            int64_t *perm);

/* RCBForceTree.cxx:525-534: the positions of a node's T pseudo-particles,
 * tdr * design + xc (T = 1 or 12). */
void hsro_pppts(int tdpts, float tdr, const float *xc, float *px, float *py,
                float *pz);

/* Particles::map1 (Particles.cxx:732-758): the step factor prefactor * tau
 * in the reference's float/double mix, and the drift x += factor * v with
 * every element rounded as x + (factor * v). */
float hsro_map1_factor(float pp, float tau, float adot, float alpha);
void hsro_drift(int64_t n, float *x, float *y, float *z, const float *vx,
                const float *vy, const float *vz, float factor);

/* Particles::cic (Particles.cxx:589-641): cloud-in-cell deposit of c =
 * gpscal^3 per particle into rho[ng0*ng1*ng2 + 1] (the last slot collects
 * the corners outside the grid). */
void hsro_cic(int64_t np, const float *x, const float *y, const float *z,
              const int32_t ng[3], float gpscal, float *rho);
/* Particles::inverse_cic (Particles.cxx:645-714), one component:
 * v += fscal * tau * CIC interpolation of grid (grid[Ng] read as 0). */
void hsro_inverse_cic(int64_t np, const float *x, const float *y, const float *z,
                      const int32_t ng[3], const float *grid, float tau, float fscal,
                      float *v);

/* Particles::map2's kick coefficient (Particles.cxx:1230-1233). */
float hsro_map2_coeff(float gpscal, double fscal, double tau,
                      double step_fraction);

/* Domain::initialize for a one-rank decomposition (Domain.cxx:34-84). */
typedef struct {
  float rL, oL_in;
  int32_t ng, ng_overload, ng_local_total;
  float oL, phys2grid_pos, grid2phys_pos, phys2grid_vel, grid2phys_vel;
} hsro_domain;
void hsro_domain_init(hsro_domain *d);

/* The refresh of MC3Extras::particleExtras on one rank, before its
 * resortParticles: copyAliveIntoVectors, ParticleExchange with every
 * neighbour this rank, copyAllFromVectors.  The ten arrays hold n particles
 * and have room for cap; returns the refreshed count (-1 if it exceeds cap,
 * the arrays then hold an unspecified mix) and the alive count in *alive. */
int64_t hsro_refresh(const hsro_domain *d, float anow, int64_t n, int64_t cap,
                     float *x, float *y, float *z, float *vx, float *vy, float *vz,
                     float *mass, float *phi, int64_t *id, uint16_t *mask,
                     int64_t *alive);

#ifdef __cplusplus
}
#endif

#endif
