// ref_harness.cpp -- TEST INFRASTRUCTURE: C entry points around the
// reference's own force-law and centre-of-mass sources, compiled where they
// lie under the reference tree (see oracle/Makefile, target `ref`).  The
// result, oracle/_ref/libhsrref.so, pins the oracle's force laws and its cm
// restatement bit-for-bit.  Nothing in the product links it.
//
// ForceLaw.cxx defines its evaluators `inline` (INLINEQ, ForceLaw.cxx:6), so
// they are only emitted in a translation unit that also uses them; the
// source is therefore included into this unit rather than linked.
#include "ForceLaw.cxx"

#include <cstdint>

extern "C" void cm(int count, float *xx, float *yy, float *zz, float *mass,
                   float *xmin, float *xmax, float *xc);  // BGQCM.c:181

extern "C" {

// BGQCM.c:181-212 (generic branch), called as RCBForceTree.cxx:785 does.
void hsrref_cm(int64_t count, float *x, float *y, float *z, float *m,
               float *xmin, float *xmax, float *xc) {
  cm((int)count, x, y, z, m, xmin, xmax, xc);
}

// ForceLawSR(FGridEvalPoly(FGrid), rsm)::f_over_r (stock POLY_ORDER 6).
float hsrref_sr_poly_f_over_r(float r2, float rsm) {
  static FGrid fg;
  FGridEvalPoly ev(&fg);
  ForceLawSR fl(&ev, rsm);
  return fl.f_over_r(r2);
}

// ForceLawSR(FGridEvalFit(FGrid), rsm)::f_over_r (production default law).
float hsrref_sr_fit_f_over_r(float r2, float rsm) {
  static FGrid fg;
  FGridEvalFit ev(&fg);
  ForceLawSR fl(&ev, rsm);
  return fl.f_over_r(r2);
}

// FGridEvalFit::eval(r2) = FGrid::fgor(sqrt(r2)): the fitted grid force/r
// that the polynomial laws approximate.
float hsrref_fgor_r2(float r2) {
  static FGrid fg;
  FGridEvalFit ev(&fg);
  return ev.eval(r2);
}

// FGridEvalPoly::eval(r2), POLY_ORDER 6.
float hsrref_poly_eval(float r2) {
  static FGrid fg;
  FGridEvalPoly ev(&fg);
  return ev.eval(r2);
}

// FGridEvalInterp(FGrid, 1024)::eval(r2) (ForceLaw.cxx:145-171).
float hsrref_interp_eval(float r2) {
  static FGrid fg;
  static FGridEvalInterp ev(&fg, 1024);
  return ev.eval(r2);
}

// ForceLawSR(FGridEvalInterp(FGrid, n), rsm)::f_over_r and FGrid's r^2
// table (ForceLaw.cxx:54-67, 145-171, 175-192) for any table size n >= 2.
float hsrref_sr_interp_f_over_r(float r2, float rsm, int n) {
  static FGrid fg;
  FGridEvalInterp ev(&fg, n);
  ForceLawSR fl(&ev, rsm);
  return fl.f_over_r(r2);
}

void hsrref_interp_table(int n, float *r2, float *f) {
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

float hsrref_newton_f_over_r(float r2) {
  ForceLawNewton fl;
  return fl.f_over_r(r2);
}

float hsrref_fgrid_rmax(void) {
  FGrid fg;
  return fg.rmax();
}

}  // extern "C"
