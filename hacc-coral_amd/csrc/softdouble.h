// softdouble.h -- exact emulation, in integer arithmetic, of the reference's
// centre-of-mass accumulation
//
//     double x = 0;  for (i ...) x += (double)(float)(w_i * xx_i);
//     xc = (float)(x / m);                         (BGQCM.c:186-211)
//
// on the GPU without FP64 instructions.  Everything here is
// __host__ __device__ so the same code is unit-tested on the CPU against
// native double arithmetic.
//
// Representation: sd_t {m, e, s} is (-1)^s * m * 2^e with m == 0 (zero) or
// 2^52 <= m < 2^53 (normalised).  The sequential sum is reproduced exactly by
// three interchangeable steps per chunk of addends:
//
//  * serial     -- sd_add_f32: one correctly rounded double addition per
//                  addend (any signs).
//  * exact      -- if every prefix sum of S + v_0 + ... + v_k is a multiple of
//                  2^t below 2^(t+53) (t = lowest set bit among S and the v's),
//                  no addition rounds and the result is the exact integer sum.
//  * fixed grid -- if S >= 0 sits in binade [2^(E+52), 2^(E+53)) and stays
//                  there, every addition rounds to the same grid 2^E, so
//                  S_k = S_{k-1} + RN_grid(v_k) except at exact half-grid
//                  ties, whose round-half-even outcome depends only on the
//                  parity of S_{k-1}/2^E.  Each addend is therefore the map
//                  n -> n + d[n & 1]; maps compose associatively, so a chunk
//                  or a whole block reduces in parallel to one map (sd_map).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SD_FN __host__ __device__ __forceinline__
#else
#define SD_FN static inline
#endif

struct sd_t {
  uint64_t m;
  int32_t e;
  int32_t s;
};

// parity-indexed increment map: n -> n + d[n & 1] (in grid units)
struct sd_map {
  uint64_t d0, d1;
  int32_t bad;  // addend too large for the grid (chunk would leave the binade)
};

#define SD_M53 (1ull << 53)
#define SD_M52 (1ull << 52)
#define SD_SAT (1ull << 62)

SD_FN int sd_clz64(uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return v ? __clzll((long long)v) : 64;
#else
  return v ? __builtin_clzll(v) : 64;
#endif
}

SD_FN int sd_ctz64(uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return v ? __ffsll((unsigned long long)v) - 1 : 64;
#else
  return v ? __builtin_ctzll(v) : 64;
#endif
}

SD_FN uint32_t sd_f2u(float f) {
  union { float f; uint32_t u; } c;
  c.f = f;
  return c.u;
}

SD_FN float sd_u2f(uint32_t u) {
  union { float f; uint32_t u; } c;
  c.u = u;
  return c.f;
}

SD_FN sd_t sd_zero() {
  sd_t r;
  r.m = 0;
  r.e = 0;
  r.s = 0;
  return r;
}

// Decompose a finite float into integer mantissa (< 2^24) and exponent, not
// normalised: v = (-1)^s * fm * 2^fe.  fm == 0 for +-0.
SD_FN void sd_split_f32(float v, uint32_t *fm, int32_t *fe, int32_t *fs) {
  uint32_t u = sd_f2u(v);
  *fs = (int32_t)(u >> 31);
  uint32_t ex = (u >> 23) & 255u, fr = u & 0x7fffffu;
  if (ex == 0) {
    *fm = fr;
    *fe = -149;
  } else {
    *fm = fr | 0x800000u;
    *fe = (int32_t)ex - 150;
  }
  if (*fm == 0) *fs = 0;
}

SD_FN sd_t sd_from_f32(float v) {
  uint32_t fm;
  int32_t fe, fs;
  sd_split_f32(v, &fm, &fe, &fs);
  sd_t r;
  if (fm == 0) return sd_zero();
  int sh = sd_clz64((uint64_t)fm) - 11;  // bring msb to bit 52
  r.m = (uint64_t)fm << sh;
  r.e = fe - sh;
  r.s = fs;
  return r;
}

// Right shift by k >= 1 with round-to-nearest-even, given an outside sticky.
SD_FN uint64_t sd_rshift_rne(uint64_t v, int k, int sticky) {
  if (k >= 64) {
    // only sticky information survives; v < 2^64 <= half when k >= 65
    uint64_t half_hi = (k == 64) ? (1ull << 63) : 0;
    if (k == 64 && (v > half_hi || (v == half_hi && sticky))) return 1;
    return 0;
  }
  uint64_t q = v >> k;
  uint64_t rem = v & ((1ull << k) - 1);
  uint64_t half = 1ull << (k - 1);
  if (rem > half || (rem == half && (sticky || (q & 1)))) ++q;
  return q;
}

// Normalise a positive magnitude v * 2^e (v < 2^64, with sticky below) to a
// 53-bit mantissa with round-to-nearest-even.
SD_FN sd_t sd_norm(uint64_t v, int32_t e, int32_t s, int sticky) {
  sd_t r;
  if (v == 0) return sd_zero();
  int p = 63 - sd_clz64(v);
  if (p > 52) {
    int k = p - 52;
    uint64_t q = sd_rshift_rne(v, k, sticky);
    e += k;
    if (q == SD_M53) {
      q = SD_M52;
      e += 1;
    }
    r.m = q;
  } else {
    r.m = v << (52 - p);
    e -= (52 - p);
  }
  r.e = e;
  r.s = s;
  return r;
}

// Correctly rounded double addition a + b (both normalised or zero).
SD_FN sd_t sd_add(sd_t a, sd_t b) {
  if (b.m == 0) return a.m == 0 ? sd_zero() : a;
  if (a.m == 0) return b;
  if (a.e < b.e || (a.e == b.e && a.m < b.m)) {
    sd_t t = a;
    a = b;
    b = t;
  }
  // now |a| >= |b| in exponent order (for equal exponents, by mantissa)
  int d = a.e - b.e;
  uint64_t A = a.m << 10;
  uint64_t Bv;
  int sticky;
  if (d >= 64) {
    Bv = 0;
    sticky = 1;
  } else {
    uint64_t Bs = b.m << 10;
    Bv = d ? (Bs >> d) : Bs;
    sticky = d ? ((Bs & ((1ull << d) - 1)) != 0) : 0;
  }
  int32_t E = a.e - 10;
  if (a.s == b.s) return sd_norm(A + Bv, E, a.s, sticky);
  // magnitudes subtract; |a| >= |b| so the sign is a's
  uint64_t S;
  if (sticky) {
    S = A - Bv - 1;
  } else {
    if (A == Bv) return sd_zero();
    S = A - Bv;
  }
  return sd_norm(S, E, a.s, sticky);
}

SD_FN sd_t sd_add_f32(sd_t a, float v) { return sd_add(a, sd_from_f32(v)); }

// -------------------------------------------------------------- fixed grid
// Map of one non-negative addend v onto grid 2^E (E = exponent of the
// accumulator's ulp).
SD_FN sd_map sd_map_of(float v, int32_t E) {
  sd_map r;
  r.bad = 0;
  uint32_t fm;
  int32_t fe, fs;
  sd_split_f32(v, &fm, &fe, &fs);
  if (fm == 0) {
    r.d0 = r.d1 = 0;
    return r;
  }
  if (fs) {
    r.bad = 1;
    r.d0 = r.d1 = 0;
    return r;
  }
  if (fe >= E) {
    int sh = fe - E;
    if (sh > 29) {
      r.bad = 1;
      r.d0 = r.d1 = 0;
      return r;
    }
    r.d0 = r.d1 = (uint64_t)fm << sh;
    return r;
  }
  int s = E - fe;
  if (s >= 25) {
    r.d0 = r.d1 = 0;
    return r;
  }
  uint64_t a = (uint64_t)(fm >> s);
  uint64_t rem = (uint64_t)fm & ((1ull << s) - 1);
  uint64_t half = 1ull << (s - 1);
  if (rem == half) {
    r.d0 = a + (a & 1);
    r.d1 = a + ((a + 1) & 1);
  } else {
    r.d0 = r.d1 = a + (rem > half ? 1 : 0);
  }
  return r;
}

SD_FN uint64_t sd_sat_add(uint64_t a, uint64_t b) {
  uint64_t c = a + b;
  return c > SD_SAT ? SD_SAT : c;
}

// f then h
SD_FN sd_map sd_map_compose(sd_map f, sd_map h) {
  sd_map r;
  r.bad = f.bad | h.bad;
  uint64_t a0 = f.d0, a1 = f.d1;
  r.d0 = sd_sat_add(a0, (a0 & 1) ? h.d1 : h.d0);
  r.d1 = sd_sat_add(a1, ((1 + a1) & 1) ? h.d1 : h.d0);
  return r;
}

SD_FN sd_map sd_map_identity() {
  sd_map r;
  r.d0 = r.d1 = 0;
  r.bad = 0;
  return r;
}

// Apply map g (built for grid E) to S.  Returns 1 and updates *S when S is
// non-negative, normalised in the binade whose ulp is 2^E and stays there.
SD_FN int sd_map_apply(sd_t *S, sd_map g, int32_t E) {
  if (g.bad || S->m == 0 || S->s || S->e != E) return 0;
  uint64_t d = (S->m & 1) ? g.d1 : g.d0;
  if (d >= SD_M53) return 0;
  uint64_t nm = S->m + d;
  if (nm >= SD_M53) return 0;
  S->m = nm;
  return 1;
}

// ------------------------------------------------------------------ exact
// lowest-set-bit exponent of a non-negative float (large for zero)
SD_FN int32_t sd_lsb_f32(float v) {
  uint32_t fm;
  int32_t fe, fs;
  sd_split_f32(v, &fm, &fe, &fs);
  if (fm == 0) return 1 << 20;
  return fe + sd_ctz64((uint64_t)fm);
}

SD_FN int32_t sd_lsb(sd_t S) {
  if (S.m == 0) return 1 << 20;
  return S.e + sd_ctz64(S.m);
}

// v as an integer multiple of 2^t; returns UINT64_MAX when v >= 2^(t+53),
// negative, or not a multiple of 2^t.  The mantissa is not normalised to an
// odd integer, so t may lie above the split exponent (1.0 = 2^23 * 2^-23 is
// a multiple of 2^0): the low mantissa bits below 2^t must then be zero.
SD_FN uint64_t sd_units_f32(float v, int32_t t) {
  uint32_t fm;
  int32_t fe, fs;
  sd_split_f32(v, &fm, &fe, &fs);
  if (fm == 0) return 0;
  if (fs) return ~0ull;
  int sh = fe - t;
  if (sh < 0) {
    const int k = -sh;
    if (k >= 32 || (fm & ((1u << k) - 1u))) return ~0ull;
    return (uint64_t)(fm >> k);
  }
  int bits = 64 - sd_clz64((uint64_t)fm);
  if (sh + bits > 53) return ~0ull;
  return (uint64_t)fm << sh;
}

SD_FN uint64_t sd_units(sd_t S, int32_t t) {
  if (S.m == 0) return 0;
  if (S.s) return ~0ull;
  int sh = S.e - t;
  // S.m has 53 significant bits, so only sh <= 0 stays below 2^53
  if (sh >= 0) {
    if (sh > 0) return ~0ull;
    return S.m;
  }
  // sh < 0: exact only if the shifted-out bits are zero
  int k = -sh;
  if (k >= 64) return ~0ull;
  if (S.m & ((1ull << k) - 1)) return ~0ull;
  return S.m >> k;
}

SD_FN sd_t sd_from_units(uint64_t u, int32_t t) {
  if (u == 0) return sd_zero();
  return sd_norm(u, t, 0, 0);
}

// ---------------------------------------------------------------- division
// (float)(a / b) with the double quotient rounded to nearest even first,
// then rounded again to float (BGQCM.c:209-211).  Normal float results only
// (coordinates of this path are far from the float subnormal range); zero
// numerator gives +-0, zero denominator gives NaN/inf like IEEE.
SD_FN float sd_div_to_f32(sd_t a, sd_t b) {
  if (b.m == 0) {
    if (a.m == 0) return sd_u2f(0x7fc00000u);
    return sd_u2f(((uint32_t)(a.s ^ b.s) << 31) | 0x7f800000u);
  }
  if (a.m == 0) return sd_u2f((uint32_t)(a.s ^ b.s) << 31);
  // quotient of mantissas in [0.5, 2): compute 56 significant bits
  uint64_t r = a.m, q = 0;
  int32_t e = a.e - b.e;
  // first bit
  int first = r >= b.m;
  if (!first) {
    r <<= 1;
    e -= 1;
  }
  // now r/b.m in [1, 2): produce 56 bits (1 integer + 55 fraction)
  for (int i = 0; i < 56; ++i) {
    q <<= 1;
    if (r >= b.m) {
      r -= b.m;
      q |= 1;
    }
    r <<= 1;
  }
  // q has 56 bits with msb at bit 55; value = q * 2^(e - 55)
  int sticky = r != 0;
  // round to 53 bits: drop 3 bits
  uint64_t q53 = sd_rshift_rne(q, 3, sticky);
  int32_t e53 = e - 55 + 3;
  if (q53 == SD_M53) {
    q53 = SD_M52;
    e53 += 1;
  }
  // double value q53 * 2^e53 -> float: drop 29 bits (no further sticky)
  uint64_t q24 = sd_rshift_rne(q53, 29, 0);
  int32_t e24 = e53 + 29;
  if (q24 == (1ull << 24)) {
    q24 = 1ull << 23;
    e24 += 1;
  }
  // float bits: value = q24 * 2^e24 with q24 in [2^23, 2^24)
  int32_t bexp = e24 + 150;
  uint32_t sign = (uint32_t)(a.s ^ b.s) << 31;
  if (bexp >= 255) return sd_u2f(sign | 0x7f800000u);
  if (bexp <= 0) {
    // subnormal result: redo the rounding at the subnormal position from
    // the double value (double rounding as the hardware conversion does)
    int k = 29 + (1 - bexp);
    uint64_t qs = sd_rshift_rne(q53, k > 63 ? 64 : k, 0);
    return sd_u2f(sign | (uint32_t)qs);
  }
  return sd_u2f(sign | ((uint32_t)bexp << 23) | ((uint32_t)q24 & 0x7fffffu));
}
