// Prints the drop-in ForceLaw hierarchy's values on an r^2 grid read from
// stdin (one float per line): ForceLawSR+Poly, ForceLawSR+Fit, fgor, poly6,
// interp(1024) -- for bitwise comparison with the reference's ForceLaw.cxx.
#include <cstdio>
#include <cstring>

#include "ForceLaw.h"

static unsigned bits(float f) { unsigned u; memcpy(&u, &f, 4); return u; }

int main() {
  FGrid fg;
  FGridEvalPoly poly(&fg);
  FGridEvalFit fit(&fg);
  FGridEvalInterp interp(&fg, 1024);
  ForceLawSR srp(&poly, 0.1f), srf(&fit, 0.1f);
  float r2;
  while (scanf("%f", &r2) == 1)
    printf("%08x %08x %08x %08x %08x\n", bits(srp.f_over_r(r2)), bits(srf.f_over_r(r2)),
           bits(fit.eval(r2)), bits(poly.eval(r2)), bits(interp.eval(r2)));
  printf("rmax %08x\n", bits(fg.rmax()));
  return 0;
}
