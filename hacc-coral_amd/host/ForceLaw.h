// ForceLaw.h -- the force-law plug-in surface of the short-range path, kept
// name-for-name with the reference (ForceLaw.h:10-127) so callers such as
// Particles.cxx:104-114,161-183 compile unchanged.
//
// On the host these classes evaluate f_over_r exactly as the reference does
// (ForceLaw.cxx); the device leaf kernel implements the laws listed in
// hacc_sr.h and RCBForceTree maps a ForceLaw object onto one of them
// (throwing for laws the device does not implement -- there is no CPU
// fallback for the tree path).
#ifndef FORCELAW_H
#define FORCELAW_H

#include <math.h>
#include <stdlib.h>

#ifndef HACC_SR_POLY_ORDER
#define HACC_SR_POLY_ORDER 6  // the reference's stock POLY_ORDER (ForceLaw.cxx:8)
#endif

class FGrid {
 public:
  FGrid();
  ~FGrid() {}
  float fgor(float r);
  void fgor_r2_interp(int nInterp, float **r2, float **f);
  float rmax() { return m_rmax; }

 protected:
  float m_b, m_c, m_d, m_e, m_f, m_g, m_h, m_l, m_rmax;
};

class FGridEval {
 public:
  FGridEval() {}
  virtual ~FGridEval() {}
  virtual float eval(float) = 0;
  virtual float r2min() = 0;
  virtual float r2max() = 0;
};

class FGridEvalFit : public FGridEval {
 public:
  FGridEvalFit(FGrid *fg);
  ~FGridEvalFit() {}
  float eval(float);
  float r2min() { return 0.0; }
  float r2max() { return m_fg->rmax() * m_fg->rmax(); }

 protected:
  FGrid *m_fg;
};

class FGridEvalPoly : public FGridEval {
 public:
  FGridEvalPoly(FGrid *fg);
  // explicit polynomial order (4, 5 or 6) instead of the compile-time default
  FGridEvalPoly(FGrid *fg, int order);
  ~FGridEvalPoly() {}
  float eval(float);
  float r2min() { return 0.0; }
  float r2max() { return m_fg->rmax() * m_fg->rmax(); }
  int order() const { return m_order; }
  const float *coefficients() const { return m_a; }

 protected:
  void init(FGrid *fg, int order);
  FGrid *m_fg;
  float m_r2min, m_r2max;
  float m_a[7];
  int m_order;
};

class FGridEvalInterp : public FGridEval {
 public:
  FGridEvalInterp(FGrid *fg, int nInterp);
  ~FGridEvalInterp();
  float eval(float);
  float r2min() { return m_r2min; }
  float r2max() { return m_r2max; }
  int nInterp() { return m_nInterp; }
  float *r2() { return m_r2; }
  float *f() { return m_f; }

 protected:
  float *m_r2;
  float *m_f;
  float m_r2min;
  float m_r2max;
  float m_dr2;
  float m_oodr2;
  int m_nInterp;
};

class ForceLaw {
 public:
  ForceLaw() {}
  virtual ~ForceLaw() {}
  virtual float f_over_r(float r2) = 0;
};

class ForceLawNewton : public ForceLaw {
 public:
  ~ForceLawNewton() {}
  float f_over_r(float r2) { return 1.0 / r2 / sqrt(r2); }
};

class ForceLawSR : public ForceLaw {
 public:
  ForceLawSR(FGridEval *fgore, float rsm);
  ~ForceLawSR() {}
  float f_over_r(float r2);
  FGridEval *grid_eval() const { return m_fgore; }
  float rsm() const { return m_rsm; }

 protected:
  float m_rsm;
  float m_rsm2;
  float m_r2min;
  float m_r2max;
  FGridEval *m_fgore;
};

// FGrid::fgor_r2_interp into caller-owned arrays (for front ends without
// the classes, e.g. Python through ctypes).
extern "C" void hsr_cxx_fgrid_table(int n, float *r2, float *f);

#endif  // FORCELAW_H
