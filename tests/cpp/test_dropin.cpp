// Drop-in demonstration and GPU test driver: the reference's caller pattern
// (Particles::map2, Particles.cxx:1311-1335) against the MI355X
// RCBForceTree.  Writes the permuted arrays and kicked velocities to a raw
// file for comparison with the oracle:
//   test_dropin <n> <box> <nd> <seed> <out.bin> [tdpts: 1 (default) or 12]
//               [law: step16 (default; the BG/Q kernel) | poly5 | poly6 | fit |
//                interp<N> (ForceLawSR over FGridEvalInterp with N points),
//                the last four through the generic ForceLaw kernel]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "RCBForceTree.h"

static uint64_t lcg(uint64_t &s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  return s >> 11;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s n box nd seed out.bin\n", argv[0]);
    return 2;
  }
  const long n = atol(argv[1]);
  const float box = (float)atof(argv[2]);
  const long nd = atol(argv[3]);
  uint64_t s = (uint64_t)atoll(argv[4]);
  std::vector<float> x(n), y(n), z(n), vx(n, 0.f), vy(n, 0.f), vz(n, 0.f), m(n, 1.f), phi(n, 0.f);
  std::vector<ID_T> id(n);
  std::vector<MASK_T> mask(n, 0);
  for (long i = 0; i < n; ++i) {
    x[i] = box * (float)((lcg(s) >> 11) * (1.0 / 9007199254740992.0 * 2048.0));
    y[i] = box * (float)((lcg(s) >> 11) * (1.0 / 9007199254740992.0 * 2048.0));
    z[i] = box * (float)((lcg(s) >> 11) * (1.0 / 9007199254740992.0 * 2048.0));
    id[i] = i;
  }
  std::vector<float> x0 = x, y0 = y, z0 = z;
  // tree box [0, box], force box [1, box - 1] (Particles.cxx:1213-1228 layout)
  float zero[3] = {0.f, 0.f, 0.f}, hi[3] = {box, box, box};
  float lo_f[3] = {1.f, 1.f, 1.f}, hi_f[3] = {box - 1.f, box - 1.f, box - 1.f};
  FGrid fg;
  const char *law = argc > 7 ? argv[7] : "step16";
  FGridEval *ge = nullptr;
  if (!strcmp(law, "step16") || !strcmp(law, "poly6")) ge = new FGridEvalPoly(&fg, 6);
  else if (!strcmp(law, "poly5")) ge = new FGridEvalPoly(&fg, 5);
  else if (!strcmp(law, "fit")) ge = new FGridEvalFit(&fg);
  else if (!strncmp(law, "interp", 6)) ge = new FGridEvalInterp(&fg, atoi(law + 6));
  else {
    fprintf(stderr, "unknown law %s\n", law);
    return 2;
  }
  ForceLawSR fl(ge, 0.1f);
  const RCBKernel kern = strcmp(law, "step16") ? RCBKernel::ForceLaw : RCBKernel::Step16;
  RCBMonopoleForceTree::setKernel(kern);
  RCBQuadrupoleForceTree::setKernel(kern);
  const float fsrrmax = fg.rmax();
  const float c = 0.5f;  // fcoeff
  const int tdpts = argc > 6 ? atoi(argv[6]) : 1;
  if (tdpts == 12) {
    RCBQuadrupoleForceTree *sft = new RCBQuadrupoleForceTree(
        zero, hi, lo_f, hi_f, n, x.data(), y.data(), z.data(), vx.data(),
        vy.data(), vz.data(), m.data(), phi.data(), id.data(), mask.data(), 1.0,
        fsrrmax, 0.1f, 0.1f, nd, 2, 128, &fl, c);
    delete sft;
  } else {
    RCBMonopoleForceTree *sft = new RCBMonopoleForceTree(
        zero, hi, lo_f, hi_f, n, x.data(), y.data(), z.data(), vx.data(),
        vy.data(), vz.data(), m.data(), phi.data(), id.data(), mask.data(), 1.0,
        fsrrmax, 0.1f, 0.1f, nd, 2, 128, &fl, c);
    delete sft;
  }
  FILE *f = fopen(argv[5], "wb");
  if (!f) return 1;
  fwrite(x0.data(), 4, n, f); fwrite(y0.data(), 4, n, f); fwrite(z0.data(), 4, n, f);
  fwrite(x.data(), 4, n, f); fwrite(y.data(), 4, n, f); fwrite(z.data(), 4, n, f);
  fwrite(vx.data(), 4, n, f); fwrite(vy.data(), 4, n, f); fwrite(vz.data(), 4, n, f);
  fwrite(id.data(), 8, n, f);
  fclose(f);
  delete ge;
  printf("dropin ok n=%ld law=%s\n", n, law);
  return 0;
}
