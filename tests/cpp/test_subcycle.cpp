// GPU test driver of the C++ sub-cycle drop-in (ParticleSubCycle): the
// reference's caller pattern, Particles::subCycle(gts) with a TimeStepper
// exposing pp(), tau(), tau2(), adot(), fscal(); and the one-rank particle
// refresh, refreshParticles.  Writes the initial and the final ten arrays to
// a raw file:
//   test_subcycle <n> <ng> <nd> <seed> <mode> <out.bin>
// mode 0: subCycle, 1: map1/map2 steps (n particles on an ng^3 grid);
// mode 2: refreshParticles(rL = ng, oL = 2.5, ng, anow = 0.8) on arrays with
// room for 3n particles (the grid is the domain's ng_local_total), the final
// arrays written at that capacity and the new count printed.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ParticleSubCycle.h"

struct Stepper {  // the accessors of TimeStepper.h:20-36 that the cycle reads
  double pp() const { return 0.5; }
  double tau() const { return 0.02; }
  double tau2() const { return 0.02; }
  double adot() const { return 1.3; }
  double fscal() const { return 1.1; }
};

static double u01(uint64_t &s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(s >> 11) * (1.0 / 9007199254740992.0);
}

int main(int argc, char **argv) {
  if (argc < 7) {
    fprintf(stderr, "usage: %s n ng nd seed mode out.bin\n", argv[0]);
    return 2;
  }
  const long n = atol(argv[1]);
  const int g = atoi(argv[2]);
  const long nd = atol(argv[3]);
  uint64_t s = (uint64_t)atoll(argv[4]);
  const int mode = atoi(argv[5]);
  hsr_domain dom{};
  dom.rL = (float)g;
  dom.oL_in = 2.5f;
  dom.ng = g;
  if (mode == 2 && hsr_domain_init(&dom) != HSR_OK) return 1;
  const int gl = mode == 2 ? dom.ng_local_total : g;  // the grid of the particles
  const long cap = mode == 2 ? 3 * n : n;
  std::vector<float> f[8];
  for (auto &a : f) a.resize(cap);
  std::vector<ID_T> id(cap);
  std::vector<MASK_T> mask(cap);
  for (long i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) f[d][i] = (float)(u01(s) * (gl + 2.0) - 1.0);
    for (int d = 3; d < 6; ++d) f[d][i] = (float)(8.0 * (u01(s) - 0.5));
    f[6][i] = (float)u01(s);  // phi
    f[7][i] = 0.5f;           // mass (map2 resets it to 1)
    id[i] = 5 * i + 1;
    mask[i] = (MASK_T)(i & 0xffff);
  }
  FILE *out = fopen(argv[6], "wb");
  if (!out) return 1;
  for (auto &a : f) fwrite(a.data(), 4, n, out);
  fwrite(id.data(), 8, n, out);
  fwrite(mask.data(), 2, n, out);
  const int ng[3] = {gl, gl, gl};
  Stepper ts;
  ParticleSubCycle P(ng, 1.0f, 1.0f, 2.0f, 5, 3.12f, 0.1f, 0.1f, nd);
  P.setParticles(n, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(),
                 f[5].data(), f[6].data(), f[7].data(), id.data(), mask.data());
  if (mode == 0)
    P.subCycle(&ts);
  else if (mode == 1)
    P.subCycleSteps(ts.pp(), ts.tau(), ts.tau2(), ts.adot(), ts.fscal());
  else
    P.refreshParticles(dom.rL, dom.oL_in, dom.ng, 0.8f, cap);
  for (auto &a : f) fwrite(a.data(), 4, cap, out);
  fwrite(id.data(), 8, cap, out);
  fwrite(mask.data(), 2, cap, out);
  fclose(out);
  printf("subcycle ok n=%ld count=%ld Np_last=%ld\n", n, (long)P.Np(), (long)P.Np_last());
  return 0;
}
