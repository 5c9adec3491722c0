// RCBForceTree.cxx -- drop-in constructor over the C ABI (see RCBForceTree.h).
#include "RCBForceTree.h"

#include <math.h>
#include <stdio.h>

#include <stdexcept>
#include <string>

#include "../../include/hacc_sr.h"

namespace {

void check(hsr_status s, const char *what) {
  if (s != HSR_OK)
    throw std::runtime_error(std::string(what) + ": " + hsr_last_error());
}

// Device law for a ForceLaw object (ForceLaw.h:95-127); *interp receives
// the FGridEvalInterp whose table the device needs.
int device_law(ForceLaw *fl, RCBKernel k, float *rsm, FGridEvalInterp **interp) {
  *interp = nullptr;
  if (k == RCBKernel::Step16) return HSR_LAW_STEP16;
  ForceLawSR *sr = dynamic_cast<ForceLawSR *>(fl);
  if (sr) {
    FGridEval *ge = sr->grid_eval();
    FGridEvalPoly *poly = dynamic_cast<FGridEvalPoly *>(ge);
    if (poly && (poly->order() == 5 || poly->order() == 6)) {
      *rsm = sr->rsm();
      return poly->order() == 5 ? HSR_LAW_SR_POLY5 : HSR_LAW_SR_POLY6;
    }
    if (dynamic_cast<FGridEvalFit *>(ge)) {
      *rsm = sr->rsm();
      return HSR_LAW_SR_FIT;
    }
    if ((*interp = dynamic_cast<FGridEvalInterp *>(ge))) {
      *rsm = sr->rsm();
      return HSR_LAW_SR_INTERP;
    }
  }
  throw std::runtime_error(
      "RCBForceTree: this ForceLaw is not implemented by the device leaf kernel "
      "(supported: ForceLawSR with FGridEvalPoly of order 5 or 6, FGridEvalFit or "
      "FGridEvalInterp, or the Step16 kernel)");
}

}  // namespace

template <int TDPTS>
RCBKernel RCBForceTree<TDPTS>::s_kernel = RCBKernel::Step16;
template <int TDPTS>
int RCBForceTree<TDPTS>::s_device = 0;

template <int TDPTS>
void RCBForceTree<TDPTS>::setKernel(RCBKernel k) {
  s_kernel = k;
}

template <int TDPTS>
void RCBForceTree<TDPTS>::setDevice(int device) {
  s_device = device;
}

template <int TDPTS>
RCBForceTree<TDPTS>::RCBForceTree(
    POSVEL_T *minLoc, POSVEL_T *maxLoc, POSVEL_T *minForceLoc,
    POSVEL_T *maxForceLoc, ID_T count, POSVEL_T *xLoc, POSVEL_T *yLoc,
    POSVEL_T *zLoc, POSVEL_T *xVel, POSVEL_T *yVel, POSVEL_T *zVel,
    POSVEL_T *mass, POSVEL_T *phiLoc, ID_T *idLoc, MASK_T *maskLoc,
    POSVEL_T avgMass, POSVEL_T fsm, POSVEL_T r, POSVEL_T oa, ID_T nd,
    ID_T ds, ID_T tmin, ForceLaw *fl, float fcoeff, POSVEL_T ppc)
    : particleCount(count), ctx(nullptr), buildTime(0) {
  (void)avgMass;  // an estimate only; unused by the reference's build too
  // RCBForceTree.cxx:395-404: no ForceLaw -> the tree's own ForceLawNewton
  // with fcoeff 1.  The BG/Q build's Step16 kernel never consults the law
  // (RCBForceTree.cxx:585-597); the generic build's nbody1 would evaluate
  // Newton's 1/r^3, which the device does not implement.
  RCBKernel k = s_kernel;
  if (!fl) {
    fcoeff = 1.0;
    if (k != RCBKernel::Step16)
      throw std::runtime_error(
          "RCBForceTree: no ForceLaw with the ForceLaw kernel means ForceLawNewton "
          "(RCBForceTree.cxx:395-404), which the device leaf kernel does not implement");
  }
  hsr_params p = {};
  p.device = s_device;
  p.tdpts = TDPTS;
  float rsm = r;
  FGridEvalInterp *interp = nullptr;
  p.law = device_law(fl, k, &rsm, &interp);
  if (interp) p.n_interp = interp->nInterp();
  p.nd = nd;
  p.ds = ds;
  p.tmin = tmin;
  p.fsrrmax = fsm;
  p.rsm = rsm;
  p.oa = oa;
  p.tan_oa = tanf(oa);  // RCBForceTree.cxx:381
  p.ppc = ppc;
  check(hsr_create(&ctx, &p), "hsr_create");
  try {
    if (interp)
      check(hsr_set_law_table(ctx, interp->r2(), interp->f(), interp->nInterp()),
            "hsr_set_law_table");
    check(hsr_build(ctx, count, xLoc, yLoc, zLoc, xVel, yVel, zVel, mass,
                    phiLoc, idLoc, maskLoc, minLoc, maxLoc, HSR_HOST_PTRS),
          "hsr_build");
    hsr_stats st;
    check(hsr_get_stats(ctx, &st), "hsr_get_stats");
    buildTime = st.build_ms * 1e-3;
    printStats(buildTime);
    check(hsr_lists(ctx, minForceLoc, maxForceLoc), "hsr_lists");
    check(hsr_kick(ctx, fcoeff), "hsr_kick");
  } catch (...) {
    hsr_destroy(ctx);
    ctx = nullptr;
    throw;
  }
}

template <int TDPTS>
RCBForceTree<TDPTS>::~RCBForceTree() {
  if (ctx) hsr_destroy(ctx);
}

// RCBForceTree.cxx:460-511 (printed when HACC_SR_PRINT_STATS is set; the
// reference prints on rank 0 only).
template <int TDPTS>
void RCBForceTree<TDPTS>::printStats(double bt) {
  if (!getenv("HACC_SR_PRINT_STATS") || !ctx) return;
  hsr_stats st;
  if (hsr_get_stats(ctx, &st) != HSR_OK) return;
  double mean = 0.0;
  if (st.leaves > 0) {
    std::vector<hsr_node> nodes(st.nodes);
    if (hsr_get_nodes(ctx, nodes.data(), st.nodes) == HSR_OK) {
      double parts = 0;
      for (size_t t = 1; t < nodes.size(); ++t)
        if (nodes[t].cl == 0 && nodes[t].cr == 0) parts += nodes[t].count;
      mean = parts / (double)st.leaves;
    }
  }
  printf("\ttree post-build statistics (local for rank 0):\n");
  printf("\t\tparticles: %.2f\n", (double)st.count);
  printf("\t\tnodes: %.2f (allocated:  %.2f)\n", (double)st.nodes, (double)st.nodes);
  printf("\t\tleaves: %.2f (empty: %.2f)\n", (double)(st.leaves + st.empty_leaves),
         (double)st.empty_leaves);
  printf("\t\tmean ppn: %.2f (max ppn: %lu)\n", mean, (unsigned long)st.max_ppn);
  printf("\t\tbuild time: %g s\n", bt);
}

template class RCBForceTree<QUADRUPOLE_TDPTS>;
template class RCBForceTree<MONOPOLE_TDPTS>;
