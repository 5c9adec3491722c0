// ParticleSubCycle.cxx -- the sub-cycle methods over the C ABI (see
// ParticleSubCycle.h).
#include "ParticleSubCycle.h"

#include <math.h>

#include <stdexcept>
#include <string>

namespace {

void check(hsr_status s, const char *what) {
  if (s != HSR_OK) throw std::runtime_error(std::string(what) + ": " + hsr_last_error());
}

}  // namespace

ParticleSubCycle::ParticleSubCycle(const int ng[DIMENSION], float alpha, float gpscal,
                                   float edge, int nsub, float fsrrmax, float rsm,
                                   float openAngle, ID_T rcbTreePPN,
                                   ID_T rcbTreeExtraLevels, int device)
    : m_alpha(alpha), m_gpscal(gpscal), m_edge(edge), m_nsub(nsub) {
  for (int d = 0; d < DIMENSION; ++d) m_ng[d] = ng[d];
  hsr_params p = {};
  p.device = device;
  p.tdpts = MONOPOLE_TDPTS;
  p.law = HSR_LAW_STEP16;
  p.nd = rcbTreePPN;
  p.ds = rcbTreeExtraLevels;
  p.tmin = 128;
  p.fsrrmax = fsrrmax;
  p.rsm = rsm;
  p.oa = openAngle;
  p.tan_oa = tanf(openAngle);  // RCBForceTree.cxx:381
  p.ppc = 0.9f;
  check(hsr_create(&m_ctx, &p), "hsr_create");
}

ParticleSubCycle::~ParticleSubCycle() {
  if (m_ctx) hsr_destroy(m_ctx);
}

void ParticleSubCycle::setParticles(ID_T Np, POSVEL_T *x, POSVEL_T *y, POSVEL_T *z,
                                    POSVEL_T *vx, POSVEL_T *vy, POSVEL_T *vz,
                                    POSVEL_T *phi, POSVEL_T *mass, ID_T *id, MASK_T *mask,
                                    bool devicePtrs) {
  m_Np = Np;
  m_x = x;
  m_y = y;
  m_z = z;
  m_vx = vx;
  m_vy = vy;
  m_vz = vz;
  m_phi = phi;
  m_mass = mass;
  m_id = id;
  m_mask = mask;
  m_flags = devicePtrs ? HSR_DEVICE_PTRS : HSR_HOST_PTRS;
}

hsr_subcycle_params ParticleSubCycle::params(double pp, double tau, double tau2, double adot,
                                             double fscal) const {
  hsr_subcycle_params sp = {};
  sp.nsub = m_nsub;
  sp.skip_stream = m_skipStreamQ ? 1 : 0;
  sp.skip_kick = m_skipKickSRQ ? 1 : 0;
  for (int d = 0; d < 3; ++d) sp.ng[d] = m_ng[d];
  sp.alpha = m_alpha;
  sp.gpscal = m_gpscal;
  sp.edge = m_edge;
  sp.pp = pp;
  sp.tau = tau;
  sp.tau2 = tau2;
  sp.adot = adot;
  sp.fscal = fscal;
  return sp;
}

void ParticleSubCycle::map1(float pp, float tau, float adot) {
  check(hsr_drift(m_ctx, m_Np, m_x, m_y, m_z, m_vx, m_vy, m_vz, pp, tau, adot, m_alpha,
                  m_flags),
        "hsr_drift");
}

void ParticleSubCycle::resortParticles() {
  int64_t inside = 0;
  const int32_t ng[3] = {m_ng[0], m_ng[1], m_ng[2]};
  check(hsr_resort(m_ctx, m_Np, ng, m_x, m_y, m_z, m_vx, m_vy, m_vz, m_mass, m_phi, m_id,
                   m_mask, &inside, m_flags),
        "hsr_resort");
  m_Np_last = inside;
}

ID_T ParticleSubCycle::refreshParticles(float rL, float oL, int ng, float anow,
                                        ID_T capacity) {
  hsr_domain dom{};
  dom.rL = rL;
  dom.oL_in = oL;
  dom.ng = ng;
  check(hsr_domain_init(&dom), "hsr_domain_init");
  for (int k = 0; k < DIMENSION; ++k)
    if (m_ng[k] != dom.ng_local_total)
      throw std::runtime_error("refreshParticles: the grid is not the domain's ng_local_total (" +
                               std::to_string(dom.ng_local_total) + ")");
  int64_t count = m_Np, inside = 0;
  check(hsr_refresh(m_ctx, &dom, anow, &count, capacity, m_x, m_y, m_z, m_vx, m_vy, m_vz,
                    m_mass, m_phi, m_id, m_mask, &inside, m_flags),
        "hsr_refresh");
  m_Np = count;
  m_Np_last = inside;
  return m_Np;
}

void ParticleSubCycle::map2(double fscal, double tau, double stepFraction) {
  const hsr_subcycle_params sp = params(0.0, tau, 0.0, 0.0, fscal);
  int64_t inside = 0;
  check(hsr_map2(m_ctx, &sp, stepFraction, m_Np, m_x, m_y, m_z, m_vx, m_vy, m_vz, m_mass,
                 m_phi, m_id, m_mask, &inside, m_flags),
        "hsr_map2");
  m_Np_last = inside;
}

void ParticleSubCycle::subCycle(double pp, double tau, double tau2, double adot,
                                double fscal) {
  const hsr_subcycle_params sp = params(pp, tau, tau2, adot, fscal);
  int64_t inside = -1;
  check(hsr_subcycle(m_ctx, &sp, m_Np, m_x, m_y, m_z, m_vx, m_vy, m_vz, m_mass, m_phi, m_id,
                     m_mask, &inside, m_flags),
        "hsr_subcycle");
  if (inside >= 0) m_Np_last = inside;
}

void ParticleSubCycle::subCycleSteps(double pp, double tau, double tau2, double adot,
                                     double fscal) {
  // Particles.cxx:1181-1198 with the single-step methods
  const double stepFraction = 1.0 / m_nsub;
  for (int step = 0; step < m_nsub; ++step) {
    if (!m_skipStreamQ) map1((float)pp, (float)(stepFraction * tau2), (float)adot);
    if (!m_skipKickSRQ) map2(fscal, tau, stepFraction);
    if (!m_skipStreamQ) map1((float)pp, (float)(stepFraction * tau2), (float)adot);
  }
}

void ParticleSubCycle::cic(float *rho) {
  const int32_t ng[3] = {m_ng[0], m_ng[1], m_ng[2]};
  check(hsr_cic(m_ctx, m_Np, m_x, m_y, m_z, ng, m_gpscal, rho, m_flags), "hsr_cic");
}

void ParticleSubCycle::inverse_cic(float tau, float fscal, int comp, const float *grad) {
  POSVEL_T *v[4] = {m_vx, m_vy, m_vz, m_phi};
  if (comp < 0 || comp > 3 || !v[comp])
    throw std::runtime_error("inverse_cic: component must be 0..3 (vx, vy, vz, phi)");
  const int32_t ng[3] = {m_ng[0], m_ng[1], m_ng[2]};
  check(hsr_inverse_cic(m_ctx, m_Np, m_x, m_y, m_z, ng, grad, tau, fscal, v[comp], m_flags),
        "hsr_inverse_cic");
}

// Particles.cxx:718-726: phi = 0, then the interpolation with tau = fscal = 1.
void ParticleSubCycle::inverse_cic_potential(const float *phiGrid) {
  if (!m_phi) throw std::runtime_error("inverse_cic_potential: no phi array");
  if (m_flags == HSR_DEVICE_PTRS)
    throw std::runtime_error("inverse_cic_potential: host arrays only (zero a device phi, "
                             "then call inverse_cic(1, 1, 3, grid))");
  for (ID_T i = 0; i < m_Np; ++i) m_phi[i] = 0.0f;
  inverse_cic(1.0f, 1.0f, 3, phiGrid);
}
