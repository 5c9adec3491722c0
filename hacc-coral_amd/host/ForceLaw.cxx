// ForceLaw.cxx -- host evaluation of the force laws (restating
// ForceLaw.cxx:23-192 of the reference: the same constants, the same float
// operation order, so f_over_r agrees bit for bit with the reference build).
#include "ForceLaw.h"

// FGrid fit constants (ForceLaw.cxx:23-33)
FGrid::FGrid()
    : m_b(0.72), m_c(0.01), m_d(0.27), m_e(0.0001), m_f(360.0), m_g(100.0),
      m_h(0.67), m_l(17.0), m_rmax(3.116326355) {}

// Fitted grid force divided by r (ForceLaw.cxx:39-51); r <= 0 takes the
// limit value c + 2 b^3 / 3.
float FGrid::fgor(float r) {
  float f0 = m_c + 2.0 / 3.0 * m_b * m_b * m_b;
  float r2 = r * r;
  float r4 = r2 * r2;
  float r6 = r4 * r2;
  float coshbr = coshf(m_b * r);
  float r3fgor = tanhf(m_b * r) - m_b * r / coshbr / coshbr +
                 m_c * r * r2 * (1.0 + m_d * r2) * expf(-1.0 * m_d * r2) +
                 m_e * r2 * (m_f * r2 + m_g * r4 + m_l * r6) * expf(-1.0 * m_h * r2);
  float rdiv = r + 1.0 * (r <= 0.0);
  float rdiv3 = rdiv * rdiv * rdiv;
  return (r3fgor / rdiv3 + (r <= 0.0) * f0) * (r <= m_rmax) * (r >= 0.0);
}

// Tabulation on a uniform r^2 grid (ForceLaw.cxx:54-67).
void FGrid::fgor_r2_interp(int nInterp, float **r2, float **f) {
  *r2 = (float *)malloc(nInterp * sizeof(float));
  *f = (float *)malloc(nInterp * sizeof(float));
  double dr2 = (m_rmax * m_rmax) / (nInterp - 1.0);
  for (int i = 0; i < nInterp; i++) {
    (*r2)[i] = i * dr2;
    (*f)[i] = fgor(sqrt(i * dr2));
  }
}

FGridEvalFit::FGridEvalFit(FGrid *fg) { m_fg = fg; }

float FGridEvalFit::eval(float r2) { return m_fg->fgor(sqrt(r2)); }

// Polynomial fits of fgor in r^2 (ForceLaw.cxx:87-131).
FGridEvalPoly::FGridEvalPoly(FGrid *fg) { init(fg, HACC_SR_POLY_ORDER); }

FGridEvalPoly::FGridEvalPoly(FGrid *fg, int order) { init(fg, order); }

void FGridEvalPoly::init(FGrid *fg, int order) {
  static const float p6[7] = {0.271431, -0.0783394, 0.0133122, -0.00159485,
                              0.000132336, -0.00000663394, 0.000000147305};
  static const float p5[7] = {0.269327, -0.0750978, 0.0114808, -0.00109313,
                              0.0000605491, -0.00000147177, 0.0};
  static const float p4[7] = {0.263729, -0.0686285, 0.00882248, -0.000592487,
                              0.0000164622, 0.0, 0.0};
  m_fg = fg;
  m_order = (order == 4 || order == 5) ? order : 6;
  const float *src = m_order == 6 ? p6 : (m_order == 5 ? p5 : p4);
  for (int k = 0; k < 7; ++k) m_a[k] = src[k];
  m_r2min = 0.0;
  m_r2max = fg->rmax() * fg->rmax();
}

float FGridEvalPoly::eval(float r2) {
  float ret = 0.0;
  ret = m_a[0] + r2 * (m_a[1] + r2 * (m_a[2] + r2 * (m_a[3] + r2 * (m_a[4] +
        r2 * (m_a[5] + r2 * m_a[6])))));
  return ret * (r2 >= m_r2min) * (r2 <= m_r2max);
}

// Linear interpolation in r^2 (ForceLaw.cxx:145-171).
FGridEvalInterp::FGridEvalInterp(FGrid *fg, int nInterp) {
  m_nInterp = nInterp;
  fg->fgor_r2_interp(m_nInterp, &m_r2, &m_f);
  m_r2min = m_r2[0];
  m_r2max = m_r2[m_nInterp - 1];
  m_dr2 = (m_r2max - m_r2min) / (m_nInterp - 1.0);
  m_oodr2 = 1.0 / m_dr2;
}

FGridEvalInterp::~FGridEvalInterp() {
  free(m_r2);
  free(m_f);
}

float FGridEvalInterp::eval(float r2) {
  int inRange, indx;
  float inRangef;
  inRange = (r2 > m_r2min) * (r2 < m_r2max);
  inRangef = 1.0 * inRange;
  indx = int((r2 - m_r2min) * m_oodr2) * inRange;
  return inRangef * (m_f[indx] + (r2 - m_r2[indx]) * m_oodr2 * (m_f[indx + 1] - m_f[indx]));
}

// Plummer-softened 1/r^3 minus the grid force (ForceLaw.cxx:175-192).
ForceLawSR::ForceLawSR(FGridEval *fgore, float rsm) {
  m_rsm = rsm;
  m_rsm2 = rsm * rsm;
  m_r2min = fgore->r2min();
  m_r2max = fgore->r2max();
  m_fgore = fgore;
}

float ForceLawSR::f_over_r(float r2) {
  float ret = powf(r2 + m_rsm2, -1.5) - m_fgore->eval(r2);
  ret *= (r2 >= m_r2min) * (r2 <= m_r2max);
  return ret;
}

// FGridEvalInterp's table for callers that hold no FGrid object (the Python
// front end): FGrid::fgor_r2_interp into caller-owned arrays of n floats.
extern "C" void hsr_cxx_fgrid_table(int n, float *r2, float *f) {
  FGrid fg;
  float *a = nullptr, *b = nullptr;
  fg.fgor_r2_interp(n, &a, &b);
  for (int i = 0; i < n; ++i) {
    r2[i] = a[i];
    f[i] = b[i];
  }
  free(a);
  free(b);
}
