// The reference's stand-alone tree accuracy scenario (ForceTreeTest.cxx:
// 112-302: "force on a particle due to a sphere of particles") over the
// MI355X RCBForceTree drop-in, without MPI and the BH tree:
//
//   forcetree_test <L> <rSphere> <nSphere> <theta> <nTrials> <S> <seed> <rcb|rcbm> [rcbN [tMin]]
//
// Each trial draws a test particle and a sphere centre uniformly in [0, L)^3
// and nSphere unit-mass particles uniformly in the ball of radius rSphere
// (drand48, as the reference), builds the tree with the test particle first
// (tree and force boxes both the particles' bounding box joined with
// [0, L]^3), kicks with fcoeff 1 through ForceLawSR(FGridEvalFit) under the
// ForceLaw kernel, and prints, as the reference does,
//   r   r^2 |v_tree|   r^2 |v_direct|   r^3 f_over_r(r^2) nSphere
// where v_direct is the direct float sum of d f_over_r(r^2) over the sphere.
// The test particle's index after the tree's permutation and its tree force
// components follow on the same line for the test harness.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ForceLaw.h"
#include "RCBForceTree.h"

int main(int argc, char **argv) {
  if (argc < 9) {
    fprintf(stderr, "usage: %s L rSphere nSphere theta nTrials S seed rcb|rcbm [rcbN [tMin]]\n",
            argv[0]);
    return 2;
  }
  const float L = (float)atof(argv[1]);
  const float rSphere = (float)atof(argv[2]);
  const int nSphere = atoi(argv[3]);
  const float theta = (float)atof(argv[4]);
  const int trials = atoi(argv[5]);
  if (argv[6][0] != 'S') {
    fprintf(stderr, "only the short-range law (S) runs on the device\n");
    return 2;
  }
  const long seed = atol(argv[7]);
  const bool mono = strcmp(argv[8], "rcbm") == 0;
  const int rcbN = argc > 9 ? atoi(argv[9]) : 12;
  const int tMin = argc > 10 ? atoi(argv[10]) : 128;
  const float rsm = 0.1f;
  const int Np = nSphere + 1;
  std::vector<float> x(Np), y(Np), z(Np), vx(Np), vy(Np), vz(Np), m(Np), phi(Np);
  std::vector<ID_T> id(Np);
  std::vector<MASK_T> mask(Np);
  FGrid fg;
  FGridEvalFit fit(&fg);
  ForceLawSR fl(&fit, rsm);
  RCBMonopoleForceTree::setKernel(RCBKernel::ForceLaw);
  RCBQuadrupoleForceTree::setKernel(RCBKernel::ForceLaw);
  const float fsrrmax = fg.rmax();
  srand48(seed);
  float zero[3] = {0.f, 0.f, 0.f}, ngl[3] = {L, L, L};
  for (int t = 0; t < trials; ++t) {
    float pp[3], ps[3];
    for (int i = 0; i < 3; ++i) {
      pp[i] = L * drand48();
      ps[i] = L * drand48();
    }
    const float dx = ps[0] - pp[0], dy = ps[1] - pp[1], dz = ps[2] - pp[2];
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float r = std::sqrt(r2);
    for (int i = 0; i < 3; ++i) {
      zero[i] = std::min(zero[i], pp[i]);
      ngl[i] = std::max(ngl[i], pp[i]);
    }
    for (int i = 0; i < Np; ++i) {
      x[i] = y[i] = z[i] = vx[i] = vy[i] = vz[i] = phi[i] = 0.f;
      m[i] = 1.f;
      id[i] = i;
      mask[i] = 0;
    }
    x[0] = pp[0];
    y[0] = pp[1];
    z[0] = pp[2];
    for (int p2 = 1; p2 < Np;) {
      const float sdx = 2.0 * rSphere * (drand48() - 0.5);
      const float sdy = 2.0 * rSphere * (drand48() - 0.5);
      const float sdz = 2.0 * rSphere * (drand48() - 0.5);
      if (sdx * sdx + sdy * sdy + sdz * sdz > rSphere * rSphere) continue;
      x[p2] = ps[0] + sdx;
      y[p2] = ps[1] + sdy;
      z[p2] = ps[2] + sdz;
      const float pt[3] = {x[p2], y[p2], z[p2]};
      for (int i = 0; i < 3; ++i) {
        zero[i] = std::min(zero[i], pt[i]);
        ngl[i] = std::max(ngl[i], pt[i]);
      }
      ++p2;
    }
    float lo[3], hi[3], flo[3], fhi[3];
    for (int i = 0; i < 3; ++i) {
      lo[i] = flo[i] = zero[i];
      hi[i] = fhi[i] = ngl[i];
    }
    if (mono) {
      RCBMonopoleForceTree tree(lo, hi, flo, fhi, Np, x.data(), y.data(), z.data(), vx.data(),
                                vy.data(), vz.data(), m.data(), phi.data(), id.data(), mask.data(),
                                1.0, fsrrmax, rsm, theta, rcbN, 2, tMin, &fl, 1.0f);
    } else {
      RCBQuadrupoleForceTree tree(lo, hi, flo, fhi, Np, x.data(), y.data(), z.data(), vx.data(),
                                  vy.data(), vz.data(), m.data(), phi.data(), id.data(),
                                  mask.data(), 1.0, fsrrmax, rsm, theta, rcbN, 2, tMin, &fl, 1.0f);
    }
    int pidx = 0;
    while (x[pidx] != pp[0]) ++pidx;
    const float f0 = std::sqrt(vx[pidx] * vx[pidx] + vy[pidx] * vy[pidx] + vz[pidx] * vz[pidx]);
    float v[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < Np; ++i) {
      if (i == pidx) continue;
      const float ex = x[i] - x[pidx], ey = y[i] - y[pidx], ez = z[i] - z[pidx];
      const float s2 = ex * ex + ey * ey + ez * ez;
      const float fr = fl.f_over_r(s2);
      v[0] += ex * fr;
      v[1] += ey * fr;
      v[2] += ez * fr;
    }
    const float f1 = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    printf("%f\t%e\t%e\t%e\t%d\t%.9e\t%.9e\t%.9e\t%.9e\t%.9e\t%.9e\n", r, r2 * f0, r2 * f1,
           r2 * r * fl.f_over_r(r2) * nSphere, pidx, vx[pidx], vy[pidx], vz[pidx], v[0], v[1], v[2]);
  }
  return 0;
}
