// ParticleSubCycle.h -- drop-in for the methods of the reference's Particles
// class around the short-range kick (Particles.h:160-175): map1 (drift,
// Particles.cxx:732-758), resortParticles (Particles.cxx:402-488), map2 with
// the RCB monopole tree (Particles.cxx:1205-1280), subCycle
// (Particles.cxx:1176-1201), the PM particle-grid steps cic and
// inverse_cic (Particles.cxx:589-725) and the one-rank particle refresh
// (MC3Extras.cxx:660-700), over the C ABI of libhaccsr.so.  The ten particle
// arrays stay with the caller (host
// or device memory); subCycle is one hsr_subcycle call, so the particles stay
// on the device for all nsub sub-steps.  The TimeStepper is any type with the
// reference's accessors pp(), tau(), tau2(), adot() and fscal()
// (TimeStepper.h:20-36), the reference's own class included.  Failures throw
// std::runtime_error carrying hsr_last_error().
#ifndef ParticleSubCycle_h
#define ParticleSubCycle_h

#include <stdint.h>

#include "../../include/hacc_sr.h"
#include "RCBForceTree.h"

class ParticleSubCycle {
 public:
  // ng: Domain::ng_local_total; alpha, gpscal, edge, nsub and the tree
  // arguments as Particles holds them (m_alpha, m_gpscal, m_edge, m_nsub,
  // m_fsrrmax, m_rsm, m_openAngle, m_rcbTreePPN, m_rcbTreeExtraLevels).
  ParticleSubCycle(const int ng[DIMENSION], float alpha, float gpscal, float edge,
                   int nsub, float fsrrmax, float rsm, float openAngle,
                   ID_T rcbTreePPN, ID_T rcbTreeExtraLevels = 2, int device = 0);
  ~ParticleSubCycle();
  ParticleSubCycle(const ParticleSubCycle &) = delete;
  ParticleSubCycle &operator=(const ParticleSubCycle &) = delete;

  // The particle arrays (m_xArr ... m_maskArr) and their count
  // (m_Np_local_total); devicePtrs: arrays are device memory.
  void setParticles(ID_T Np, POSVEL_T *x, POSVEL_T *y, POSVEL_T *z, POSVEL_T *vx,
                    POSVEL_T *vy, POSVEL_T *vz, POSVEL_T *phi, POSVEL_T *mass,
                    ID_T *id, MASK_T *mask, bool devicePtrs = false);

  void map1(float pp, float tau, float adot);
  void resortParticles();
  void map2(double fscal, double tau, double stepFraction = 1.0);
  template <class TimeStepper>
  void map2(TimeStepper *ts, double stepFraction = 1.0) {
    map2(ts->fscal(), ts->tau(), stepFraction);
  }
  template <class TimeStepper>
  void subCycle(TimeStepper *gts) {
    subCycle(gts->pp(), gts->tau(), gts->tau2(), gts->adot(), gts->fscal());
  }
  void subCycle(double pp, double tau, double tau2, double adot, double fscal);
  // the same cycle through map1 / map2 / map1 per sub-step
  void subCycleSteps(double pp, double tau, double tau2, double adot, double fscal);

  // Particles::cic / inverse_cic (Particles.cxx:589-714) with the grid
  // passed in (the reference's m_field): ng[0]*ng[1]*ng[2] + 1 floats, the
  // last the slot for corners outside the grid.  comp 0, 1, 2 kick vx, vy,
  // vz; 3 adds to phi (inverse_cic_potential zeroes phi first).
  void cic(float *rho);
  void inverse_cic(float tau, float fscal, int comp, const float *grad);
  void inverse_cic_potential(const float *phiGrid);

  void setSkip(bool skipStream, bool skipKickSR) {
    m_skipStreamQ = skipStream;
    m_skipKickSRQ = skipKickSR;
  }
  // The particle refresh of MC3Extras::particleExtras on one rank
  // (MC3Extras.cxx:330-338, 660-700) for the domain of the deck's rL, oL and
  // ng (Domain::initialize): alive particles kept in order, the overload
  // copies of ParticleExchange appended (every neighbour is this rank), unit
  // masses, grid units, resortParticles.  The arrays given to setParticles
  // must have room for `capacity` particles; the grid of this object must be
  // the domain's ng_local_total.  Returns the new particle count, which the
  // object keeps (m_Np_local_total).
  ID_T refreshParticles(float rL, float oL, int ng, float anow, ID_T capacity);
  // m_Np_last: particles inside the grid after the last resort
  ID_T Np_last() const { return m_Np_last; }
  ID_T Np() const { return m_Np; }

 private:
  hsr_subcycle_params params(double pp, double tau, double tau2, double adot,
                             double fscal) const;
  int m_ng[DIMENSION];
  float m_alpha, m_gpscal, m_edge;
  int m_nsub;
  bool m_skipStreamQ = false, m_skipKickSRQ = false;
  ID_T m_Np = 0, m_Np_last = -1;
  POSVEL_T *m_x = nullptr, *m_y = nullptr, *m_z = nullptr;
  POSVEL_T *m_vx = nullptr, *m_vy = nullptr, *m_vz = nullptr;
  POSVEL_T *m_phi = nullptr, *m_mass = nullptr;
  ID_T *m_id = nullptr;
  MASK_T *m_mask = nullptr;
  int m_flags = HSR_HOST_PTRS;
  hsr_ctx *m_ctx = nullptr;
};

#endif  // ParticleSubCycle_h
