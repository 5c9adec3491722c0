// Host unit test of hacc-coral_amd/csrc/softdouble.h against native IEEE
// double arithmetic (x86-64 SSE2, round to nearest even).
#include "../../hacc-coral_amd/csrc/softdouble.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

static uint64_t dbits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

static double sd_to_double(sd_t s) {
  if (s.m == 0) return 0.0;
  double v = std::ldexp((double)s.m, s.e);
  return s.s ? -v : v;
}

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { if (fails < 20) { printf(__VA_ARGS__); printf("\n"); } ++fails; } } while (0)

static float rand_float(std::mt19937_64 &g, int mode) {
  std::uniform_real_distribution<double> u(0.0, 1.0);
  switch (mode) {
  case 0: return (float)(u(g) * 262.4);                       // box coordinates
  case 1: return (float)std::ldexp(u(g), (int)(g() % 60) - 40); // wide range
  case 2: return (float)((u(g) - 0.5) * std::ldexp(1.0, (int)(g() % 30) - 10));
  case 3: { uint32_t b = (uint32_t)g() & 0x7fffffffu;          // any finite
            if (((b >> 23) & 255) == 255) b &= 0x7f7fffffu;
            float f; memcpy(&f, &b, 4); return (g() & 1) ? -f : f; }
  case 4: return (float)((g() % 4096) * 0.25);                  // many ties
  case 6: return (float)(1 + g() % 8);                          // unit-like masses
  default: return (float)(u(g) * 4.0);
  }
}

int main() {
  std::mt19937_64 g(12345);
  // 1. serial addition == native double accumulation
  for (int trial = 0; trial < 4000; ++trial) {
    int mode = trial % 6;
    int n = 1 + (int)(g() % 3000);
    double acc = 0.0;
    sd_t s = sd_zero();
    for (int i = 0; i < n; ++i) {
      float v = rand_float(g, mode);
      if (mode == 3 && !std::isfinite((double)v)) v = 1.0f;
      acc += (double)v;
      s = sd_add_f32(s, v);
    }
    CHECK(dbits(acc) == dbits(sd_to_double(s)) || (acc == 0 && s.m == 0),
          "serial mismatch mode %d n %d: %.17g vs %.17g", mode, n, acc, sd_to_double(s));
  }
  // 2. chunk paths (exact, fixed grid) agree with serial when they apply
  long applied_exact = 0, applied_map = 0;
  for (int trial = 0; trial < 20000; ++trial) {
    int mode = (trial % 4 == 0) ? 4 : (trial % 4 == 1 ? 0 : (trial % 4 == 2 ? 5 : 6));
    sd_t S = sd_zero();
    int pre = (int)(g() % 5000);
    for (int i = 0; i < pre; ++i) S = sd_add_f32(S, rand_float(g, mode));
    float v[64];
    for (int k = 0; k < 64; ++k) v[k] = rand_float(g, mode);
    sd_t ser = S;
    for (int k = 0; k < 64; ++k) ser = sd_add_f32(ser, v[k]);
    // exact
    int32_t t = sd_lsb(S);
    for (int k = 0; k < 64; ++k) { int32_t l = sd_lsb_f32(v[k]); if (l < t) t = l; }
    uint64_t tot = sd_units(S, t);
    bool ok = tot != ~0ull;
    for (int k = 0; k < 64 && ok; ++k) { uint64_t u = sd_units_f32(v[k], t); if (u == ~0ull) ok = false; else tot += u; }
    if (ok && tot < SD_M53) {
      sd_t r = sd_from_units(tot, t);
      ++applied_exact;
      CHECK(r.m == ser.m && r.e == ser.e, "exact path mismatch");
    }
    // fixed grid
    if (S.m != 0) {
      sd_map f = sd_map_identity();
      for (int k = 0; k < 64; ++k) f = sd_map_compose(f, sd_map_of(v[k], S.e));
      // also check associativity: compose in a tree order
      sd_map lv[64];
      for (int k = 0; k < 64; ++k) lv[k] = sd_map_of(v[k], S.e);
      for (int w = 1; w < 64; w *= 2)
        for (int k = 0; k + w < 64; k += 2 * w) lv[k] = sd_map_compose(lv[k], lv[k + w]);
      CHECK(lv[0].d0 == f.d0 && lv[0].d1 == f.d1 && lv[0].bad == f.bad, "compose not associative");
      sd_t r = S;
      if (sd_map_apply(&r, f, S.e)) {
        ++applied_map;
        CHECK(r.m == ser.m && r.e == ser.e, "map path mismatch trial %d", trial);
      }
    }
  }
  // 3. division then double->float rounding
  for (int trial = 0; trial < 200000; ++trial) {
    int mode = trial % 6;
    double a = 0, b = 0;
    sd_t A = sd_zero(), B = sd_zero();
    int n = 1 + (int)(g() % 50);
    for (int i = 0; i < n; ++i) {
      float v = rand_float(g, mode == 3 ? 1 : mode);
      float w = (trial % 7 == 0) ? 1.0f : (float)(0.5 + (g() % 1000) / 500.0);
      float vw = w * v;
      a += (double)vw; A = sd_add_f32(A, vw);
      b += (double)w;  B = sd_add_f32(B, w);
    }
    float ref = (float)(a / b);
    float got = sd_div_to_f32(A, B);
    uint32_t r1, r2; memcpy(&r1, &ref, 4); memcpy(&r2, &got, 4);
    CHECK(r1 == r2, "div mismatch %.9g vs %.9g", ref, got);
  }
  printf("applied exact %ld map %ld fails %d\n", applied_exact, applied_map, fails);
  return fails ? 1 : 0;
}
