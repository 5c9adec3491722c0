// wave.h -- wave64 reduction helpers for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "softdouble.h"

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Exact, order-independent reductions (integer min / max / sum, float min /
// max) by DPP: an inclusive scan within each row of 16 lanes (row_shr 1, 2,
// 4, 8), then row 0's total into row 1 and row 2's into row 3 (row_bcast
// 15), then rows 0..1's into rows 2..3 (row_bcast 31); lane 63 holds the
// wave's result, which readlane broadcasts.  Disabled source lanes give the
// identity.
template <int CTRL, int RMASK>
__device__ __forceinline__ int dpp_i32(int v, int ident) {
  return __builtin_amdgcn_update_dpp(ident, v, CTRL, RMASK, 0xf, false);
}

#define HSR_DPP_REDUCE(v, ident, OP)                        \
  do {                                                      \
    v = OP(v, dpp_i32<0x111, 0xf>(v, ident));               \
    v = OP(v, dpp_i32<0x112, 0xf>(v, ident));               \
    v = OP(v, dpp_i32<0x114, 0xf>(v, ident));               \
    v = OP(v, dpp_i32<0x118, 0xf>(v, ident));               \
    v = OP(v, dpp_i32<0x142, 0xa>(v, ident));               \
    v = OP(v, dpp_i32<0x143, 0xc>(v, ident));               \
  } while (0)

__device__ __forceinline__ int dpp_op_min(int a, int b) { return min(a, b); }
__device__ __forceinline__ int dpp_op_max(int a, int b) { return max(a, b); }
__device__ __forceinline__ int dpp_op_fmin(int a, int b) {
  return __float_as_int(fminf(__int_as_float(a), __int_as_float(b)));
}
__device__ __forceinline__ int dpp_op_fmax(int a, int b) {
  return __float_as_int(fmaxf(__int_as_float(a), __int_as_float(b)));
}

__device__ __forceinline__ int wave_min_i32(int v) {
  HSR_DPP_REDUCE(v, INT_MAX, dpp_op_min);
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int wave_max_i32(int v) {
  HSR_DPP_REDUCE(v, INT_MIN, dpp_op_max);
  return __builtin_amdgcn_readlane(v, 63);
}

template <int CTRL, int RMASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, RMASK, 0xf, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, RMASK, 0xf, false);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  v += dpp_u64<0x111, 0xf>(v);
  v += dpp_u64<0x112, 0xf>(v);
  v += dpp_u64<0x114, 0xf>(v);
  v += dpp_u64<0x118, 0xf>(v);
  v += dpp_u64<0x142, 0xa>(v);
  v += dpp_u64<0x143, 0xc>(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
  return (int64_t)wave_sum_u64((uint64_t)v);
}

__device__ __forceinline__ float wave_sum_f32(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ float wave_min_f32(float v) {
  int b = __float_as_int(v);
  HSR_DPP_REDUCE(b, __float_as_int(INFINITY), dpp_op_fmin);
  return __int_as_float(__builtin_amdgcn_readlane(b, 63));
}

__device__ __forceinline__ float wave_max_f32(float v) {
  int b = __float_as_int(v);
  HSR_DPP_REDUCE(b, __float_as_int(-INFINITY), dpp_op_fmax);
  return __int_as_float(__builtin_amdgcn_readlane(b, 63));
}

// In-order composition of the lanes' maps (lane 0's element first); the
// result is returned in every lane.
__device__ __forceinline__ sd_map wave_compose_maps(sd_map m) {
  int l = lane_id();
  for (int o = 1; o < 64; o <<= 1) {
    sd_map h;
    h.d0 = (uint64_t)__shfl_down((unsigned long long)m.d0, o);
    h.d1 = (uint64_t)__shfl_down((unsigned long long)m.d1, o);
    h.bad = __shfl_down(m.bad, o);
    if ((l & (2 * o - 1)) == 0) m = sd_map_compose(m, h);
  }
  m.d0 = (uint64_t)__shfl((unsigned long long)m.d0, 0);
  m.d1 = (uint64_t)__shfl((unsigned long long)m.d1, 0);
  m.bad = __shfl(m.bad, 0);
  return m;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int readlane_i32(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

// Sequential correctly rounded double sum S += buf[0] + ... + buf[len-1]
// (addends staged in LDS, every lane sees them), reproduced exactly and with
// all 64 lanes working.  When no addition of the rest can round, the exact
// integer sum is the answer.  Otherwise lane l composes the fixed-grid map of a contiguous
// run of the remaining addends for S's binade, an ordered inclusive scan of
// the lane maps gives every prefix's map, and the longest prefix that keeps
// S in its binade is applied at once.  The run that leaves the binade is
// then added serially and the rest resumes in the new binade, so a block
// costs one scan per binade it crosses.
__device__ __forceinline__ void sd_block(sd_t &S, const float *buf, int len) {
  const int l = lane_id();
  int start = 0;
  while (start < len) {  // uniform
    const int per = (len - start + 63) / 64;
    const int lo = start + l * per, hi = min(lo + per, len);
    if (!S.s) {
      // no rounding anywhere in the rest (every partial sum a multiple of the
      // finest lsb below 2^53 of them): the sum is the exact integer sum
      int32_t t = S.m ? sd_lsb(S) : (1 << 20);
      for (int e = lo; e < hi; ++e) t = min(t, sd_lsb_f32(buf[e]));
      t = wave_min_i32(t);
      uint64_t u = 0;
      int bad = 0;
      for (int e = lo; e < hi; ++e) {
        const uint64_t uu = sd_units_f32(buf[e], t);
        if (uu == ~0ull) bad = 1;
        else u += uu;
      }
      const uint64_t su = sd_units(S, t);
      if (!__any(bad) && su != ~0ull) {
        const uint64_t tot = wave_sum_u64(u) + su;
        if (tot < SD_M53) {
          S = sd_from_units(tot, t);
          return;
        }
      }
    }
    if (S.m == 0 || S.s) {  // from zero (or a negative sum): one addend exactly
      S = sd_add_f32(S, buf[start]);
      ++start;
      continue;
    }
    const int32_t E = S.e;
    sd_map m = sd_map_identity();
    for (int e = lo; e < hi; ++e) m = sd_map_compose(m, sd_map_of(buf[e], E));
    for (int o = 1; o < 64; o <<= 1) {  // ordered inclusive scan
      sd_map h;
      h.d0 = (uint64_t)__shfl_up((unsigned long long)m.d0, o);
      h.d1 = (uint64_t)__shfl_up((unsigned long long)m.d1, o);
      h.bad = __shfl_up(m.bad, o);
      if (l >= o) m = sd_map_compose(h, m);
    }
    sd_t S2 = S;
    const bool ok = sd_map_apply(&S2, m, E) != 0;
    const uint64_t bal = (uint64_t)__ballot(ok);
    const int j = (~bal == 0ull) ? 64 : (int)__builtin_ctzll(~bal);  // ok lanes: a prefix
    if (j == 64) {
      S.m = readlane_u64(S2.m, 63);
      S.e = readlane_i32(S2.e, 63);
      S.s = readlane_i32(S2.s, 63);
      return;
    }
    if (j > 0) {
      S.m = readlane_u64(S2.m, j - 1);
      S.e = readlane_i32(S2.e, j - 1);
      S.s = readlane_i32(S2.s, j - 1);
    }
    const int jlo = start + j * per, jhi = min(jlo + per, len);
    for (int e = jlo; e < jhi; ++e) S = sd_add_f32(S, buf[e]);
    start = jhi;
  }
}

// Accumulate one 64-addend chunk (lane k holds addend k; lanes >= nvalid
// hold 0) into the wave-uniform exact double S, reproducing k sequential
// correctly rounded double additions.  The exact-sum test is cheapest and
// covers most chunks; the fixed-grid map covers chunks that round inside one
// binade; the serial path covers the rest (binade crossings, signs).
__device__ __forceinline__ void sd_chunk(sd_t &S, float v, int nvalid) {
  {
    int t = wave_min_i32(min(sd_lsb_f32(v), sd_lsb(S)));
    uint64_t u = sd_units_f32(v, t);
    int bad = __any(u == ~0ull);
    uint64_t su = sd_units(S, t);
    if (!bad && su != ~0ull) {
      uint64_t tot = wave_sum_u64(u) + su;
      if (tot < SD_M53) {
        S = sd_from_units(tot, t);
        return;
      }
    }
  }
  if (S.m != 0 && S.s == 0) {
    sd_map mp = wave_compose_maps(sd_map_of(v, S.e));
    sd_t S2 = S;
    if (sd_map_apply(&S2, mp, S.e)) {
      S = S2;
      return;
    }
  }
  for (int k = 0; k < nvalid; ++k) {
    float vk = __shfl(v, k);
    S = sd_add_f32(S, vk);
  }
}

