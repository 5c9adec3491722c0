// api.hip -- the C ABI of libhaccsr.so (include/hacc_sr.h).
//
// The entry points stage the caller's ten particle arrays (host or device
// memory), run the device tree build, list build and leaf kernel on the
// context's stream, and write the permuted arrays back in place exactly as
// RCBForceTree's constructor does (RCBForceTree.cxx:336-450).
#include "hsr_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

namespace {

thread_local std::string g_err;

// BGQStep16.c:40-45,167 and ForceLaw.cxx:109-114 (POLY_ORDER 5)
const float kP5[6] = {0.269327f, -0.0750978f, 0.0114808f,
                      -0.00109313f, 0.0000605491f, -0.00000147177f};
// ForceLaw.cxx:98-104 (POLY_ORDER 6)
const float kP6[7] = {0.271431f, -0.0783394f, 0.0133122f, -0.00159485f,
                      0.000132336f, -0.00000663394f, 0.000000147305f};
// FGrid::m_rmax, ForceLaw.cxx:32
const float kRmax = 3.116326355f;

// The device law for a ForceLaw choice.  ForceLawSR zeroes r2 > rmax^2
// (ForceLaw.cxx:189) and nbody1 zeroes r2 >= fsrrmax^2
// (RCBForceTree.cxx:613); both become one strict cutoff rc2.
hsr_status make_law(const hsr_params &p, ForceLawDev &L) {
  memset(&L, 0, sizeof(L));
  float fsr2 = p.fsrrmax * p.fsrrmax;
  L.eps2 = p.rsm * p.rsm;
  switch (p.law) {
  case HSR_LAW_STEP16:
    for (int i = 0; i < 6; ++i) L.c[i] = kP5[i];
    L.order = 5;
    L.rc2 = fsr2;
    L.guard = L.eps2 > 0.f ? 0 : 1;
    break;
  case HSR_LAW_SR_POLY5:
  case HSR_LAW_SR_POLY6: {
    if (p.law == HSR_LAW_SR_POLY5) {
      for (int i = 0; i < 6; ++i) L.c[i] = kP5[i];
      L.order = 5;
    } else {
      for (int i = 0; i < 7; ++i) L.c[i] = kP6[i];
      L.order = 6;
    }
    float r2max = kRmax * kRmax;
    // r2 <= r2max  <=>  r2 < nextafter(r2max, +inf)
    float incl = std::nextafter(r2max, INFINITY);
    L.rc2 = std::min(fsr2, incl);
    // r2 == 0 pairs have dx = 0 and contribute nothing when eps2 > 0
    L.guard = L.eps2 > 0.f ? 0 : 1;
    break;
  }
  case HSR_LAW_SR_FIT: {
    // ForceLawSR masks r2 in [0, rmax^2] with FGridEvalFit's r2max()
    // (ForceLaw.h:47, rmax()*rmax() in float); nbody1 masks r2 < fsrrmax^2
    L.kind = LAWK_FIT;
    float incl = std::nextafter(kRmax * kRmax, INFINITY);
    L.rc2 = std::min(fsr2, incl);
    L.guard = L.eps2 > 0.f ? 0 : 1;
    break;
  }
  case HSR_LAW_SR_INTERP: {
    // the cutoff follows the table's last r2 once hsr_set_law_table gives it
    if (p.n_interp < 2) {
      g_err = "HSR_LAW_SR_INTERP needs params.n_interp >= 2";
      return HSR_ERR_ARG;
    }
    L.kind = LAWK_INTERP;
    L.n_interp = p.n_interp;
    L.rc2 = fsr2;
    L.guard = L.eps2 > 0.f ? 0 : 1;
    break;
  }
  default:
    g_err = "unsupported force law (device path implements STEP16, SR_POLY5, SR_POLY6, "
            "SR_FIT and SR_INTERP)";
    return HSR_ERR_UNSUPPORTED;
  }
  return HSR_OK;
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hsr_status check_ctx(hsr_ctx *c) {
  if (!c) {
    g_err = "null context";
    return HSR_ERR_ARG;
  }
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) {
    g_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return HSR_ERR_HIP;
  }
  return HSR_OK;
}

struct Gather6 {
  const float *in[4];
  float *out[4];
  const int64_t *id_in;
  int64_t *id_out;
  const uint16_t *mask_in;
  uint16_t *mask_out;
};

// out[k] = in[P[k]] for the present arrays of vx, vy, vz, phi, id, mask in
// one pass over the permutation.
__global__ void k_gather_rest(int64_t n, const int32_t *__restrict__ P, Gather6 g) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t s = P[k];
  // the staging buffers are distinct from the arrays read: every value is
  // loaded before any is stored, so the gathers are in flight together
  float v[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) v[a] = g.in[a] ? g.in[a][s] : 0.f;
  const int64_t idv = g.id_in ? g.id_in[s] : 0;
  const uint16_t mv = g.mask_in ? g.mask_in[s] : (uint16_t)0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
    if (g.in[a]) g.out[a][k] = v[a];
  if (g.id_in) g.id_out[k] = idv;
  if (g.mask_in) g.mask_out[k] = mv;
}

// Permute the ten arrays (device memory) in place with the build's final
// permutation.  x, y, z and mass come from the sorted copies the build made
// (the same gather); the other six are gathered in one pass into staging
// and copied back.
hsr_status permute_in_place(hsr_ctx *c, int64_t n, float *f[8], int64_t *id,
                            uint16_t *mask) {
  hipStream_t st = c->stream;
  const int32_t *P = dptr<int32_t>(c->tree.pfin);
  const float *sorted[4] = {dptr<float>(c->tree.xs), dptr<float>(c->tree.ys),
                            dptr<float>(c->tree.zs), dptr<float>(c->tree.ms)};
  const int own[4] = {0, 1, 2, 6};
  for (int k = 0; k < 4; ++k) {
    if (k == 3 && c->tree.unit) continue;  // all ones: any permutation leaves them
    HSR_HIP(hipMemcpyAsync(f[own[k]], sorted[k], sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  }
  const int rest[4] = {3, 4, 5, 7};
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  HSR_HIP(dbuf_reserve(c->s.x1, (sizeof(int64_t) + 4 * sizeof(float) + sizeof(uint16_t)) * nn + 256));
  char *sp = (char *)c->s.x1.p;
  Gather6 g{};
  if (id) {  // 8-byte array first keeps every staging array aligned
    g.id_in = id;
    g.id_out = (int64_t *)sp;
    sp += sizeof(int64_t) * nn;
  }
  for (int a = 0; a < 4; ++a)
    if (f[rest[a]]) {
      g.in[a] = f[rest[a]];
      g.out[a] = (float *)sp;
      sp += sizeof(float) * nn;
    }
  if (mask) {
    g.mask_in = mask;
    g.mask_out = (uint16_t *)sp;
  }
  if (g.id_in || g.mask_in || g.in[0] || g.in[1] || g.in[2] || g.in[3]) {
    k_gather_rest<<<nblk(n, 256), 256, 0, st>>>(n, P, g);
    HSR_HIP(hipGetLastError());
  }
  for (int a = 0; a < 4; ++a)
    if (g.in[a])
      HSR_HIP(hipMemcpyAsync(f[rest[a]], g.out[a], sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  if (id) HSR_HIP(hipMemcpyAsync(id, g.id_out, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, st));
  if (mask) HSR_HIP(hipMemcpyAsync(mask, g.mask_out, sizeof(uint16_t) * n, hipMemcpyDeviceToDevice, st));
  return HSR_OK;
}

hsr_status build_common(hsr_ctx *c, int64_t nboxes, const int64_t *box_off,
                        float *x, float *y, float *z, float *vx, float *vy,
                        float *vz, float *mass, float *phi, int64_t *id,
                        uint16_t *mask, const float *min_loc,
                        const float *max_loc, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (nboxes < 1 || !box_off || !x || !y || !z || !mass || !min_loc || !max_loc) {
    g_err = "hsr_build: null argument";
    return HSR_ERR_ARG;
  }
  int64_t n = box_off[nboxes];
  for (int64_t b = 0; b < nboxes; ++b)
    if (box_off[b + 1] < box_off[b]) {
      g_err = "hsr_build_many: offsets must be non-decreasing";
      return HSR_ERR_ARG;
    }
  if (box_off[0] != 0 || n < 0 || n >= (int64_t)INT32_MAX) {
    g_err = "hsr_build: particle count must be in [0, 2^31)";
    return HSR_ERR_ARG;
  }
  hipStream_t st = c->stream;
  c->built = c->listed = false;
  float *f[8] = {x, y, z, vx, vy, vz, mass, phi};
  float *df[8];
  int64_t *did = id;
  uint16_t *dmask = mask;
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  if (ptr_flags & HSR_DEVICE_PTRS) {
    for (int k = 0; k < 8; ++k) df[k] = f[k];
  } else {
    for (int k = 0; k < 8; ++k) {
      df[k] = nullptr;
      if (!f[k]) continue;
      HSR_HIP(dbuf_reserve(c->s.host_in[k], sizeof(float) * nn));
      df[k] = dptr<float>(c->s.host_in[k]);
      if (n) HSR_HIP(hipMemcpyAsync(df[k], f[k], sizeof(float) * n, hipMemcpyHostToDevice, st));
    }
    did = nullptr;
    dmask = nullptr;
    if (id) {
      HSR_HIP(dbuf_reserve(c->s.host_in[8], sizeof(int64_t) * nn));
      did = dptr<int64_t>(c->s.host_in[8]);
      if (n) HSR_HIP(hipMemcpyAsync(did, id, sizeof(int64_t) * n, hipMemcpyHostToDevice, st));
    }
    if (mask) {
      HSR_HIP(dbuf_reserve(c->s.host_in[9], sizeof(uint16_t) * nn));
      dmask = dptr<uint16_t>(c->s.host_in[9]);
      if (n) HSR_HIP(hipMemcpyAsync(dmask, mask, sizeof(uint16_t) * n, hipMemcpyHostToDevice, st));
    }
  }
  // the build's time covers the tree and the permutation of the caller's
  // arrays (the reference's constructor does both)
  HSR_HIP(hipEventRecord(c->ev[0], st));
  s = tree_build_device(c, nboxes, box_off, min_loc, max_loc, df[0], df[1], df[2], df[6]);
  if (s) return s;
  if (n) {
    s = permute_in_place(c, n, df, did, dmask);
    if (s) return s;
  }
  HSR_HIP(hipEventRecord(c->ev[1], st));
  if (!(ptr_flags & HSR_DEVICE_PTRS) && n) {
    for (int k = 0; k < 8; ++k)
      if (f[k]) HSR_HIP(hipMemcpyAsync(f[k], df[k], sizeof(float) * n, hipMemcpyDeviceToHost, st));
    if (id) HSR_HIP(hipMemcpyAsync(id, did, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
    if (mask) HSR_HIP(hipMemcpyAsync(mask, dmask, sizeof(uint16_t) * n, hipMemcpyDeviceToHost, st));
  }
  HSR_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->build_ms = ms;
  c->user_v[0] = vx;
  c->user_v[1] = vy;
  c->user_v[2] = vz;
  c->user_mass = mass;
  c->user_flags = ptr_flags;
  // device-side velocity arrays for hsr_kick
  if (!(ptr_flags & HSR_DEVICE_PTRS)) {
    c->user_dv[0] = df[3];
    c->user_dv[1] = df[4];
    c->user_dv[2] = df[5];
  } else {
    c->user_dv[0] = vx;
    c->user_dv[1] = vy;
    c->user_dv[2] = vz;
  }
  c->built = true;
  return HSR_OK;
}

// Particles::map1's step factor prefactor * tau in the reference's float
// arithmetic (Particles.cxx:746-747): powf of pp by the double exponent
// 1 + 1/alpha rounded to float, then 1.0 / (alpha * adot * pf) in double,
// stored as float.
float map1_factor(float pp, float tau, float adot, float alpha) {
  const float pf = powf(pp, (float)(1.0 + 1.0 / (double)alpha));
  const float prefactor = (float)(1.0 / (double)(alpha * adot * pf));
  return prefactor * tau;
}

// Particles::map2's kick coefficient (Particles.cxx:1230-1233): float
// gpscal^3 and pi = 4 atanf(1), the rest in double, stored as float.
float map2_coeff(float gpscal, double fscal, double tau, double stepFraction) {
  const float divscal = gpscal * gpscal * gpscal;
  const float pi = (float)(4.0 * (double)atanf(1.0f));
  return (float)((double)divscal / 4.0 / (double)pi * fscal * tau * stepFraction);
}

hsr_status sub_events(hsr_ctx *c) {
  for (int k = 0; k < 6; ++k)
    if (!c->sub.ev[k]) HSR_HIP(hipEventCreate(&c->sub.ev[k]));
  return HSR_OK;
}

hsr_status check_count_grid(const char *who, int64_t count, const int32_t *ng) {
  if (count < 0 || count >= (int64_t)INT32_MAX) {
    g_err = std::string(who) + ": particle count must be in [0, 2^31)";
    return HSR_ERR_ARG;
  }
  if (!ng || ng[0] < 1 || ng[1] < 1 || ng[2] < 1 || ng[0] > (1 << 24) || ng[1] > (1 << 24) ||
      ng[2] > (1 << 24) ||
      (int64_t)ng[0] * ng[1] * ng[2] >= (int64_t)INT32_MAX) {
    g_err = std::string(who) + ": ng must be positive with ng[0]*ng[1]*ng[2] < 2^31 - 1";
    return HSR_ERR_ARG;
  }
  return HSR_OK;
}

// Copy the present arrays among the ten into device buffers buf[0..9].
hsr_status stage_in(hsr_ctx *c, DBuf *buf, int64_t n, float *const f[8], const int64_t *id,
                    const uint16_t *mask, float *df[8], int64_t **did, uint16_t **dmask) {
  hipStream_t st = c->stream;
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  for (int a = 0; a < 8; ++a) {
    df[a] = nullptr;
    if (!f[a]) continue;
    HSR_HIP(dbuf_reserve(buf[a], sizeof(float) * nn));
    df[a] = dptr<float>(buf[a]);
    if (n) HSR_HIP(hipMemcpyAsync(df[a], f[a], sizeof(float) * n, hipMemcpyHostToDevice, st));
  }
  *did = nullptr;
  *dmask = nullptr;
  if (id) {
    HSR_HIP(dbuf_reserve(buf[8], sizeof(int64_t) * nn));
    *did = dptr<int64_t>(buf[8]);
    if (n) HSR_HIP(hipMemcpyAsync(*did, id, sizeof(int64_t) * n, hipMemcpyHostToDevice, st));
  }
  if (mask) {
    HSR_HIP(dbuf_reserve(buf[9], sizeof(uint16_t) * nn));
    *dmask = dptr<uint16_t>(buf[9]);
    if (n) HSR_HIP(hipMemcpyAsync(*dmask, mask, sizeof(uint16_t) * n, hipMemcpyHostToDevice, st));
  }
  return HSR_OK;
}

hsr_status stage_out(hsr_ctx *c, int64_t n, float *const f[8], int64_t *id, uint16_t *mask,
                     float *const df[8], const int64_t *did, const uint16_t *dmask) {
  if (!n) return HSR_OK;
  hipStream_t st = c->stream;
  for (int a = 0; a < 8; ++a)
    if (f[a]) HSR_HIP(hipMemcpyAsync(f[a], df[a], sizeof(float) * n, hipMemcpyDeviceToHost, st));
  if (id) HSR_HIP(hipMemcpyAsync(id, did, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
  if (mask) HSR_HIP(hipMemcpyAsync(mask, dmask, sizeof(uint16_t) * n, hipMemcpyDeviceToHost, st));
  return HSR_OK;
}

}  // namespace

void hsr_set_error(const std::string &msg) { g_err = msg; }

hipError_t dbuf_reserve(DBuf &b, size_t bytes) {
  if (bytes <= b.bytes && b.p) return hipSuccess;
  // a buffer that grows again takes a quarter more than asked, so slowly
  // growing sizes (list lengths from step to step) settle on one allocation
  // and its address (which a captured force phase holds)
  size_t want = std::max<size_t>(bytes, 256);
  if (b.p) {
    want = std::max(want, b.bytes + b.bytes / 4);
    hipError_t e = hipFree(b.p);
    if (e != hipSuccess) return e;
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess && want > bytes) {  // no room for the slack: the size asked
    (void)hipGetLastError();
    want = std::max<size_t>(bytes, 256);
    e = hipMalloc(&b.p, want);
  }
  if (e != hipSuccess) {
    b.p = nullptr;
    return e;
  }
  b.bytes = want;
  return hipSuccess;
}

void dbuf_free(DBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

extern "C" {

const char *hsr_last_error(void) { return g_err.c_str(); }

const char *hsr_version(void) { return "hacc-sr-mi355x 0.1 (gfx950)"; }

hsr_status hsr_create(hsr_ctx **out, const hsr_params *p) {
  if (!out || !p) {
    g_err = "hsr_create: null argument";
    return HSR_ERR_ARG;
  }
  *out = nullptr;
  if (p->tdpts != 1 && p->tdpts != 12) {
    g_err = "hsr_create: TDPTS must be 1 (RCBMonopoleForceTree) or 12 (RCBQuadrupoleForceTree)";
    return HSR_ERR_UNSUPPORTED;
  }
  if (p->nd < 1) {
    g_err = "hsr_create: nd must be >= 1";
    return HSR_ERR_ARG;
  }
  if (!(p->fsrrmax > 0.f) || !(p->rsm >= 0.f) || !(p->ppc > 0.f)) {
    g_err = "hsr_create: fsrrmax > 0, rsm >= 0, ppc > 0 required";
    return HSR_ERR_ARG;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    g_err = "hsr_create: no HIP device available (there is no CPU path)";
    return HSR_ERR_HIP;
  }
  if (p->device < 0 || p->device >= ndev) {
    g_err = "hsr_create: device ordinal out of range";
    return HSR_ERR_ARG;
  }
  hsr_ctx *c = new hsr_ctx();
  c->p = *p;
  c->device = p->device;
  hsr_status s = make_law(*p, c->law);
  if (s) {
    delete c;
    return s;
  }
  memset(&c->tuning, 0, sizeof(c->tuning));
  if (hipSetDevice(c->device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    g_err = "hsr_create: stream creation failed";
    delete c;
    return HSR_ERR_HIP;
  }
  // the allocation-free entry points' scratch address on this device
  if ((s = noalloc_scratch_init(c->device))) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return s;
  }
  for (int k = 0; k < 8; ++k) (void)hipEventCreate(&c->ev[k]);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess)
    c->num_cus = prop.multiProcessorCount;
  *out = c;
  return HSR_OK;
}

hsr_status hsr_set_law_table(hsr_ctx *c, const float *r2, const float *f, int64_t n) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (c->law.kind != LAWK_INTERP) {
    g_err = "hsr_set_law_table: the context's law is not HSR_LAW_SR_INTERP";
    return HSR_ERR_STATE;
  }
  if (!r2 || !f || n != c->law.n_interp) {
    g_err = "hsr_set_law_table: null table or n != params.n_interp";
    return HSR_ERR_ARG;
  }
  for (int64_t i = 1; i < n; ++i)
    if (!(r2[i] > r2[i - 1]) || !std::isfinite(f[i]) || !std::isfinite(f[i - 1])) {
      g_err = "hsr_set_law_table: r2 must ascend and f be finite";
      return HSR_ERR_ARG;
    }
  // FGridEvalInterp's derived range (ForceLaw.cxx:145-153): float fields
  // assigned from double expressions
  const float r2min = r2[0], r2max = r2[n - 1];
  const float dr2 = (r2max - r2min) / (n - 1.0);
  const float oodr2 = 1.0 / dr2;
  std::vector<float4> bins((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const float fi = f[i], fn = i + 1 < n ? f[i + 1] : f[i];
    bins[(size_t)i] = make_float4(r2[i], fi, fn - fi, 0.f);
  }
  HSR_HIP(dbuf_reserve(c->law_bins, sizeof(float4) * (size_t)n));
  HSR_HIP(hipMemcpy(c->law_bins.p, bins.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice));
  c->law.bins = dptr<float4>(c->law_bins);
  c->law.t_r2min = r2min;
  c->law.t_r2max = r2max;
  c->law.t_oodr2 = oodr2;
  // ForceLawSR's mask r2 in [r2min, r2max] (ForceLaw.cxx:184-190) with
  // r2min = r2[0] = 0, and nbody1's r2 < fsrrmax^2
  const float fsr2 = c->p.fsrrmax * c->p.fsrrmax;
  c->law.rc2 = std::min(fsr2, std::nextafter(r2max, INFINITY));
  c->graph_gen = -1;  // a captured force phase holds the old law arguments
  return HSR_OK;
}

hsr_status hsr_destroy(hsr_ctx *c) {
  if (!c) return HSR_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  DBuf *bufs[] = {&c->tree.bnodes, &c->tree.canon_nodes, &c->tree.pfin, &c->tree.pj, &c->tree.ppm12,
                  &c->tree.xs, &c->tree.ys, &c->tree.zs, &c->tree.ms,
                  &c->lists.leaves, &c->lists.seg_ptr, &c->lists.seg, &c->lists.ent_ptr,
                  &c->lists.jidx, &c->lists.items, &c->lists.stack, &c->lists.walk, &c->lists.segcache,
                  &c->lists.item_keys,
                  &c->lists.plan, &c->lists.part, &c->s.qctr,
                  &c->adj.fw, &c->adj.fw_items, &c->adj.tr, &c->adj.tr_items, &c->adj.seg,
                  &c->adj.tj, &c->adj.par, &c->adj.counts, &c->adj.qctr, &c->adj.part, &c->adj.H,
                  &c->adj.W, &c->law_bins, &c->f64_at, &c->f64_g,
                  &c->s.a, &c->s.b, &c->s.c, &c->s.d, &c->s.e, &c->s.f, &c->s.g, &c->s.h,
                  &c->s.cub, &c->s.cms, &c->s.cmfail, &c->s.cmx, &c->s.x1, &c->s.y1, &c->s.z1, &c->s.m1, &c->s.p1,
                  &c->s.x0, &c->s.y0, &c->s.z0, &c->s.m0, &c->s.p0, &c->s.blocks, &c->s.rights,
                  &c->s.pblocks, &c->s.cnode, &c->s.pnode, &c->s.pmask, &c->s.counters, &c->dev_v, &c->force};
  if (c->adj.graph_exec) (void)hipGraphExecDestroy(c->adj.graph_exec);
  if (c->adj.graph) (void)hipGraphDestroy(c->adj.graph);
  for (DBuf *b : bufs) dbuf_free(*b);
  for (int k = 0; k < 10; ++k) dbuf_free(c->s.host_in[k]);
  dbuf_free(c->sub.keys);
  dbuf_free(c->sub.stage);
  dbuf_free(c->sub.xcnt);
  dbuf_free(c->sub.xstage);
  dbuf_free(c->sub.xout);
  for (int k = 0; k < 10; ++k) dbuf_free(c->sub.arr[k]);
  for (int k = 0; k < 6; ++k)
    if (c->sub.ev[k]) (void)hipEventDestroy(c->sub.ev[k]);
  for (int k = 0; k < 8; ++k) (void)hipEventDestroy(c->ev[k]);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return HSR_OK;
}

hsr_status hsr_get_stream(hsr_ctx *c, void **stream) {
  if (!c || !stream) {
    g_err = "hsr_get_stream: null argument";
    return HSR_ERR_ARG;
  }
  *stream = (void *)c->stream;
  return HSR_OK;
}

hsr_status hsr_build(hsr_ctx *c, int64_t count, float *x, float *y, float *z,
                     float *vx, float *vy, float *vz, float *mass, float *phi,
                     int64_t *id, uint16_t *mask, const float min_loc[3],
                     const float max_loc[3], int ptr_flags) {
  if (count < 0) {
    g_err = "hsr_build: negative count";
    return HSR_ERR_ARG;
  }
  int64_t off[2] = {0, count};
  return build_common(c, 1, off, x, y, z, vx, vy, vz, mass, phi, id, mask,
                      min_loc, max_loc, ptr_flags);
}

hsr_status hsr_build_many(hsr_ctx *c, int64_t nboxes, const int64_t *offsets,
                          float *x, float *y, float *z, float *vx, float *vy,
                          float *vz, float *mass, float *phi, int64_t *id,
                          uint16_t *mask, const float *min_loc,
                          const float *max_loc, int ptr_flags) {
  return build_common(c, nboxes, offsets, x, y, z, vx, vy, vz, mass, phi, id,
                      mask, min_loc, max_loc, ptr_flags);
}

hsr_status hsr_lists(hsr_ctx *c, const float min_force[3],
                     const float max_force[3]) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->built) {
    g_err = "hsr_lists: no tree (call hsr_build first)";
    return HSR_ERR_STATE;
  }
  if (c->tree.nboxes != 1) {
    g_err = "hsr_lists: context holds several boxes; use hsr_lists_many";
    return HSR_ERR_STATE;
  }
  if (!min_force || !max_force) {
    g_err = "hsr_lists: null force box";
    return HSR_ERR_ARG;
  }
  // the old lists and plan are gone from here on: until the new ones are
  // complete, every call that needs them fails with HSR_ERR_STATE
  c->listed = false;
  c->lists_gen++;
  HSR_HIP(hipEventRecord(c->ev[2], c->stream));
  s = lists_build_device(c, min_force, max_force);
  if (s) return s;
  s = force_plan(c);
  if (s) return s;
  c->lists_gen++;
  HSR_HIP(hipEventRecord(c->ev[3], c->stream));
  HSR_HIP(hipEventSynchronize(c->ev[3]));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
  c->lists_ms = ms;
  c->listed = true;
  return HSR_OK;
}

hsr_status hsr_lists_many(hsr_ctx *c, const float *min_force,
                          const float *max_force) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->built) {
    g_err = "hsr_lists_many: no tree";
    return HSR_ERR_STATE;
  }
  if (!min_force || !max_force) {
    g_err = "hsr_lists_many: null force boxes";
    return HSR_ERR_ARG;
  }
  // the old lists and plan are gone from here on: until the new ones are
  // complete, every call that needs them fails with HSR_ERR_STATE
  c->listed = false;
  c->lists_gen++;
  HSR_HIP(hipEventRecord(c->ev[2], c->stream));
  s = lists_build_device(c, min_force, max_force);
  if (s) return s;
  s = force_plan(c);
  if (s) return s;
  c->lists_gen++;
  HSR_HIP(hipEventRecord(c->ev[3], c->stream));
  HSR_HIP(hipEventSynchronize(c->ev[3]));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
  c->lists_ms = ms;
  c->listed = true;
  return HSR_OK;
}

// Pairs the last force phase evaluated (its kernel's counter, qctr[1]).
static hipError_t read_evaluated(hsr_ctx *c) {
  c->evaluated_pairs = c->unmasked_pairs = 0;
  if (c->lists.nleaves == 0 || c->lists.nitems == 0 || !c->s.qctr.p) return hipSuccess;
  hipError_t e = hipMemcpyAsync(&c->evaluated_pairs, dptr<unsigned long long>(c->s.qctr) + 1,
                                sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(&c->unmasked_pairs, dptr<unsigned long long>(c->s.qctr) + 2,
                        sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
}

hsr_status hsr_forces(hsr_ctx *c, float *ax, float *ay, float *az,
                      int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->listed) {
    g_err = "hsr_forces: no lists (call hsr_lists first)";
    return HSR_ERR_STATE;
  }
  if (!ax || !ay || !az) {
    g_err = "hsr_forces: null output";
    return HSR_ERR_ARG;
  }
  int64_t n = c->tree.n;
  float *fx = ax, *fy = ay, *fz = az;
  if (!(ptr_flags & HSR_DEVICE_PTRS)) {
    HSR_HIP(dbuf_reserve(c->force, sizeof(float) * 3 * std::max<int64_t>(n, 1)));
    fx = dptr<float>(c->force);
    fy = fx + n;
    fz = fy + n;
  }
  HSR_HIP(hipEventRecord(c->ev[4], c->stream));
  s = force_run(c, fx, fy, fz);
  if (s) return s;
  HSR_HIP(hipEventRecord(c->ev[5], c->stream));
  HSR_HIP(read_evaluated(c));
  if (!(ptr_flags & HSR_DEVICE_PTRS) && n) {
    HSR_HIP(hipMemcpyAsync(ax, fx, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HSR_HIP(hipMemcpyAsync(ay, fy, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HSR_HIP(hipMemcpyAsync(az, fz, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  }
  HSR_HIP(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
  c->force_ms = ms;
  c->kernel_ms = 0;
  if (c->kernel_timed) {
    HSR_HIP(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
    c->kernel_ms = ms;
  }
  return HSR_OK;
}

hsr_status hsr_kick(hsr_ctx *c, float fcoeff) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->listed) {
    g_err = "hsr_kick: no lists (call hsr_lists first)";
    return HSR_ERR_STATE;
  }
  if (!c->user_dv[0] || !c->user_dv[1] || !c->user_dv[2]) {
    g_err = "hsr_kick: the build was given no velocity arrays";
    return HSR_ERR_STATE;
  }
  int64_t n = c->tree.n;
  HSR_HIP(hipEventRecord(c->ev[4], c->stream));
  s = kick_run(c, fcoeff, c->user_dv[0], c->user_dv[1], c->user_dv[2], dptr<float>(c->tree.ms));
  if (s) return s;
  HSR_HIP(hipEventRecord(c->ev[5], c->stream));
  HSR_HIP(read_evaluated(c));
  if (!(c->user_flags & HSR_DEVICE_PTRS) && n) {
    for (int k = 0; k < 3; ++k)
      HSR_HIP(hipMemcpyAsync(c->user_v[k], c->user_dv[k], sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  }
  HSR_HIP(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
  c->force_ms = ms;
  c->kernel_ms = 0;
  if (c->kernel_timed) {
    HSR_HIP(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
    c->kernel_ms = ms;
  }
  return HSR_OK;
}

hsr_status hsr_adjoint(hsr_ctx *c, const float *lx, const float *ly,
                       const float *lz, float *gx, float *gy, float *gz,
                       int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->listed) {
    g_err = "hsr_adjoint: no lists";
    return HSR_ERR_STATE;
  }
  if (!lx || !ly || !lz || !gx || !gy || !gz) {
    g_err = "hsr_adjoint: null argument";
    return HSR_ERR_ARG;
  }
  if (c->p.tdpts != 1) {
    g_err = "hsr_adjoint: the adjoint covers the monopole tree (TDPTS = 1)";
    return HSR_ERR_UNSUPPORTED;
  }
  const int64_t n = c->tree.n;
  const float *dl[3] = {lx, ly, lz};
  float *dg[3] = {gx, gy, gz};
  if (!(ptr_flags & HSR_DEVICE_PTRS)) {
    HSR_HIP(dbuf_reserve(c->s.host_in[0], sizeof(float) * 6 * std::max<int64_t>(n, 1)));
    float *b = dptr<float>(c->s.host_in[0]);
    for (int k = 0; k < 3; ++k) {
      if (n) HSR_HIP(hipMemcpyAsync(b + k * n, dl[k], sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
      dl[k] = b + k * n;
      dg[k] = b + (3 + k) * n;
    }
  }
  s = adjoint_run(c, dl[0], dl[1], dl[2], dg[0], dg[1], dg[2]);
  if (s) return s;
  if (!(ptr_flags & HSR_DEVICE_PTRS) && n) {
    HSR_HIP(hipMemcpyAsync(gx, dg[0], sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HSR_HIP(hipMemcpyAsync(gy, dg[1], sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HSR_HIP(hipMemcpyAsync(gz, dg[2], sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  }
  HSR_HIP(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
  c->kernel_ms = ms;
  c->adjoint_ms = ms;
  return HSR_OK;
}

hsr_status hsr_forces_f64(hsr_ctx *c, double *ax, double *ay, double *az,
                          double *leaf_mom, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->listed) {
    g_err = "hsr_forces_f64: no lists (call hsr_lists first)";
    return HSR_ERR_STATE;
  }
  if (!ax || !ay || !az) {
    g_err = "hsr_forces_f64: null output";
    return HSR_ERR_ARG;
  }
  const int64_t n = c->tree.n, nl = c->lists.nleaves;
  double *d[4] = {ax, ay, az, leaf_mom};
  DBuf &buf = c->s.host_in[1];
  if (!(ptr_flags & HSR_DEVICE_PTRS)) {
    HSR_HIP(dbuf_reserve(buf, sizeof(double) * (3 * (size_t)std::max<int64_t>(n, 1) + 3 * (size_t)(nl + 1))));
    double *b = dptr<double>(buf);
    d[0] = b;
    d[1] = b + n;
    d[2] = b + 2 * n;
    d[3] = leaf_mom ? b + 3 * n : nullptr;
  }
  s = forces_f64_run(c, d[0], d[1], d[2], d[3]);
  if (s) return s;
  if (!(ptr_flags & HSR_DEVICE_PTRS)) {
    if (n)
      for (int k = 0; k < 3; ++k)
        HSR_HIP(hipMemcpyAsync(k == 0 ? ax : (k == 1 ? ay : az), d[k], sizeof(double) * n,
                               hipMemcpyDeviceToHost, c->stream));
    if (leaf_mom && nl)
      HSR_HIP(hipMemcpyAsync(leaf_mom, d[3], sizeof(double) * 3 * nl, hipMemcpyDeviceToHost, c->stream));
  }
  HSR_HIP(hipStreamSynchronize(c->stream));
  return HSR_OK;
}

hsr_status hsr_forces_f64_at(hsr_ctx *c, const double *x, const double *y, const double *z,
                             double *ax, double *ay, double *az) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!x || !y || !z || !ax || !ay || !az) {
    g_err = "hsr_forces_f64_at: null argument";
    return HSR_ERR_ARG;
  }
  if ((s = forces_f64_at_run(c, x, y, z, ax, ay, az))) return s;
  HSR_HIP(hipStreamSynchronize(c->stream));
  return HSR_OK;
}

hsr_status hsr_adjoint_f64_at(hsr_ctx *c, const double *x, const double *y, const double *z,
                              const double *lx, const double *ly, const double *lz, double *gx,
                              double *gy, double *gz) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!x || !y || !z || !lx || !ly || !lz || !gx || !gy || !gz) {
    g_err = "hsr_adjoint_f64_at: null argument";
    return HSR_ERR_ARG;
  }
  const int64_t n = c->tree.n, ng = n + c->tree.nnodes + 1;
  HSR_HIP(dbuf_reserve(c->f64_g, sizeof(double) * 3 * (size_t)ng));
  double *g = dptr<double>(c->f64_g);
  if ((s = adjoint_f64_at_run(c, x, y, z, lx, ly, lz, g, g + ng, g + 2 * ng))) return s;
  double *out[3] = {gx, gy, gz};
  if (n)
    for (int k = 0; k < 3; ++k)
      HSR_HIP(hipMemcpyAsync(out[k], g + k * ng, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  HSR_HIP(hipStreamSynchronize(c->stream));
  return HSR_OK;
}

hsr_status hsr_leaf_momentum(hsr_ctx *c, const float *ax, const float *ay,
                             const float *az, double *leaf_mom, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->listed) {
    g_err = "hsr_leaf_momentum: no lists (call hsr_lists first)";
    return HSR_ERR_STATE;
  }
  if (!ax || !ay || !az || !leaf_mom) {
    g_err = "hsr_leaf_momentum: null argument";
    return HSR_ERR_ARG;
  }
  const int64_t n = c->tree.n, nl = c->lists.nleaves;
  const float *a[3] = {ax, ay, az};
  double *mom = leaf_mom;
  DBuf &buf = c->s.host_in[2];
  if (!(ptr_flags & HSR_DEVICE_PTRS)) {
    HSR_HIP(dbuf_reserve(buf, sizeof(float) * 3 * (size_t)std::max<int64_t>(n, 1) + sizeof(double) * 3 * (size_t)(nl + 2)));
    float *b = dptr<float>(buf);
    for (int k = 0; k < 3; ++k) {
      if (n) HSR_HIP(hipMemcpyAsync(b + k * n, a[k], sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
      a[k] = b + k * n;
    }
    mom = reinterpret_cast<double *>(b + 3 * std::max<int64_t>(n, 1) + (n % 2 ? 1 : 0));
  }
  s = leaf_momentum_run(c, a[0], a[1], a[2], mom);
  if (s) return s;
  if (!(ptr_flags & HSR_DEVICE_PTRS) && nl)
    HSR_HIP(hipMemcpyAsync(leaf_mom, mom, sizeof(double) * 3 * nl, hipMemcpyDeviceToHost, c->stream));
  HSR_HIP(hipStreamSynchronize(c->stream));
  return HSR_OK;
}

hsr_status hsr_get_stats(hsr_ctx *c, hsr_stats *st) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!st) {
    g_err = "hsr_get_stats: null output";
    return HSR_ERR_ARG;
  }
  memset(st, 0, sizeof(*st));
  st->count = c->tree.n;
  st->nodes = c->tree.nnodes;
  st->max_depth = c->tree.max_depth;
  if (c->built && c->tree.nnodes > 0) {
    std::vector<hsr_node> nodes(c->tree.nnodes);
    HSR_HIP(hipMemcpy(nodes.data(), c->tree.canon_nodes.p, sizeof(hsr_node) * c->tree.nnodes,
                      hipMemcpyDeviceToHost));
    // printStats (RCBForceTree.cxx:460-511) counts from node 1
    for (size_t t = 1; t < nodes.size(); ++t) {
      if (nodes[t].cl == 0 && nodes[t].cr == 0) {
        if (nodes[t].count > 0) {
          st->leaves++;
          st->max_ppn = std::max<int64_t>(st->max_ppn, nodes[t].count);
        } else {
          st->empty_leaves++;
        }
      }
    }
  }
  st->listed_leaves = c->lists.nleaves;
  st->segments = c->lists.nsegs;
  st->entries = c->lists.nentries;
  st->pairs = c->lists.pairs;
  st->evaluated_pairs = c->evaluated_pairs;
  st->unmasked_pairs = c->unmasked_pairs;
  st->graph_captures = c->graph_captures;
  st->graph_replays = c->graph_launches - c->graph_captures;
  st->build_ms = c->build_ms;
  st->lists_ms = c->lists_ms;
  st->force_ms = c->force_ms;
  st->kernel_ms = c->kernel_ms;
  st->adjoint_ms = c->adjoint_ms;
  st->drift_ms = c->sub.drift_ms;
  st->resort_ms = c->sub.resort_ms;
  st->subcycle_ms = c->sub.subcycle_ms;
  st->drift_factor = c->sub.drift_factor;
  st->kick_coeff = c->sub.kick_coeff;
  st->cic_ms = c->sub.cic_ms;
  st->inverse_cic_ms = c->sub.inverse_cic_ms;
  st->refresh_ms = c->sub.refresh_ms;
  st->refresh_alive = c->sub.refresh_alive;
  st->refresh_dead = c->sub.refresh_dead;
  return HSR_OK;
}

hsr_status hsr_get_pseudo_masses(hsr_ctx *c, float *ppm, int64_t cap) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->built) {
    g_err = "hsr_get_pseudo_masses: no tree";
    return HSR_ERR_STATE;
  }
  const int64_t nn = c->tree.nnodes, T = c->p.tdpts;
  if (!ppm || cap < nn * T) {
    g_err = "hsr_get_pseudo_masses: output holds fewer than nodes * TDPTS floats";
    return HSR_ERR_ARG;
  }
  if (nn == 0) return HSR_OK;
  if (T == 12) {
    HSR_HIP(hipMemcpy(ppm, c->tree.ppm12.p, sizeof(float) * 12 * nn, hipMemcpyDeviceToHost));
  } else {
    std::vector<hsr_node> nodes(nn);
    HSR_HIP(hipMemcpy(nodes.data(), c->tree.canon_nodes.p, sizeof(hsr_node) * nn, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < nn; ++k) ppm[k] = nodes[k].ppm[0];
  }
  return HSR_OK;
}

hsr_status hsr_get_nodes(hsr_ctx *c, hsr_node *out, int64_t cap) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->built) {
    g_err = "hsr_get_nodes: no tree";
    return HSR_ERR_STATE;
  }
  if (!out || cap < c->tree.nnodes) {
    g_err = "hsr_get_nodes: output too small";
    return HSR_ERR_ARG;
  }
  if (c->tree.nnodes)
    HSR_HIP(hipMemcpy(out, c->tree.canon_nodes.p, sizeof(hsr_node) * c->tree.nnodes,
                      hipMemcpyDeviceToHost));
  return HSR_OK;
}

hsr_status hsr_get_perm(hsr_ctx *c, int64_t *perm, int64_t cap) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->built) {
    g_err = "hsr_get_perm: no tree";
    return HSR_ERR_STATE;
  }
  int64_t n = c->tree.n;
  if (!perm || cap < n) {
    g_err = "hsr_get_perm: output too small";
    return HSR_ERR_ARG;
  }
  std::vector<int32_t> p(n);
  if (n) HSR_HIP(hipMemcpy(p.data(), c->tree.pfin.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < n; ++k) perm[k] = p[k];
  return HSR_OK;
}

hsr_status hsr_get_lists(hsr_ctx *c, int64_t *leaves, int64_t *seg_ptr,
                         int64_t *seg_a, int64_t *seg_b, int64_t *entries,
                         int64_t cap_leaves, int64_t cap_segs) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!c->listed) {
    g_err = "hsr_get_lists: no lists";
    return HSR_ERR_STATE;
  }
  ListState &L = c->lists;
  if (cap_leaves < L.nleaves || cap_segs < L.nsegs) {
    g_err = "hsr_get_lists: output too small";
    return HSR_ERR_ARG;
  }
  int64_t nl = L.nleaves, ns = L.nsegs;
  std::vector<int32_t> lv(nl);
  std::vector<int64_t> ep(nl + 1);
  if (nl) {
    HSR_HIP(hipMemcpy(lv.data(), L.leaves.p, sizeof(int32_t) * nl, hipMemcpyDeviceToHost));
    HSR_HIP(hipMemcpy(seg_ptr, L.seg_ptr.p, sizeof(int64_t) * (nl + 1), hipMemcpyDeviceToHost));
    HSR_HIP(hipMemcpy(ep.data(), L.ent_ptr.p, sizeof(int64_t) * (nl + 1), hipMemcpyDeviceToHost));
  } else if (seg_ptr) {
    seg_ptr[0] = 0;
  }
  if (ns) {
    const int64_t *sa = dptr<int64_t>(L.seg);
    HSR_HIP(hipMemcpy(seg_a, sa, sizeof(int64_t) * ns, hipMemcpyDeviceToHost));
    HSR_HIP(hipMemcpy(seg_b, sa + ns + 1, sizeof(int64_t) * ns, hipMemcpyDeviceToHost));
  }
  for (int64_t k = 0; k < nl; ++k) {
    leaves[k] = lv[k];
    if (entries) entries[k] = ep[k + 1] - ep[k];
  }
  return HSR_OK;
}

hsr_status hsr_set_tuning(hsr_ctx *c, const hsr_tuning *t) {
  if (!c || !t) {
    g_err = "hsr_set_tuning: null argument";
    return HSR_ERR_ARG;
  }
  int r, u, tile;
  force_shape(*t, &r, &u, &tile);
  if (t->i_per_lane < 0 || t->unroll < 0 || t->tile < 0 ||
      !force_shape_built(c->law.kind, c->law.guard != 0, r, u, tile)) {
    g_err = "hsr_set_tuning: (i_per_lane, unroll, tile) = (" + std::to_string(r) + ", " +
            std::to_string(u) + ", " + std::to_string(tile) + ") is not built for this law; "
            "built: " + force_shapes_built() + " (the polynomial laws with rsm > 0), the "
            "first for every law";
    return HSR_ERR_ARG;
  }
  if (t->force_pass != 0 && t->force_pass != 32 * r && t->force_pass != 64 * r) {
    g_err = "hsr_set_tuning: force_pass must be 0, " + std::to_string(32 * r) + " or " +
            std::to_string(64 * r);
    return HSR_ERR_ARG;
  }
  if ((t->chunk != 0 && (t->chunk < tile || t->chunk > (1 << 30))) || t->waves_per_cu < 0 ||
      t->waves_per_cu > 256) {
    g_err = "hsr_set_tuning: chunk must be 0 or lie in [" + std::to_string(tile) +
            ", 2^30], waves_per_cu in [0, 256]";
    return HSR_ERR_ARG;
  }
  if (t->max_plan_bytes < 0) {
    g_err = "hsr_set_tuning: max_plan_bytes must be >= 0";
    return HSR_ERR_ARG;
  }
  if (t->chain_hops < 0 || t->chain_hops > 4096) {
    g_err = "hsr_set_tuning: chain_hops must lie in [0, 4096]";
    return HSR_ERR_ARG;
  }
  const bool replan = t->tree_order_targets != c->tuning.tree_order_targets ||
                      t->chunk != c->tuning.chunk;
  const hsr_tuning old = c->tuning;
  c->tuning = *t;
  c->graph_gen = -1;  // a captured force phase holds the old settings
  if (replan && c->listed) {
    // The target order and the work items are part of the force plan.  A
    // failed re-plan leaves the plan half written: the lists count as gone
    // until a plan with the old tuning is made again.
    c->listed = false;
    c->lists_gen++;
    hsr_status s = force_plan(c);
    if (s) {
      const std::string why = g_err;
      c->tuning = old;
      if (force_plan(c) == HSR_OK) {
        c->lists_gen++;
        c->listed = true;
      }
      g_err = why;
      return s;
    }
    c->lists_gen++;
    c->listed = true;
  }
  return HSR_OK;
}

hsr_status hsr_get_tuning(hsr_ctx *c, hsr_tuning *t) {
  if (!c || !t) {
    g_err = "hsr_get_tuning: null argument";
    return HSR_ERR_ARG;
  }
  *t = c->tuning;
  return HSR_OK;
}


hsr_status hsr_drift(hsr_ctx *c, int64_t count, float *x, float *y, float *z,
                     const float *vx, const float *vy, const float *vz, float pp,
                     float tau, float adot, float alpha, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (count < 0 || !x || !y || !z || !vx || !vy || !vz) {
    g_err = "hsr_drift: null array or negative count";
    return HSR_ERR_ARG;
  }
  if ((s = sub_events(c))) return s;
  const float k = map1_factor(pp, tau, adot, alpha);
  c->sub.drift_factor = k;
  float *d[6] = {x, y, z, (float *)vx, (float *)vy, (float *)vz};
  const bool host = !(ptr_flags & HSR_DEVICE_PTRS);
  hipStream_t st = c->stream;
  if (host) {
    for (int a = 0; a < 6; ++a) {
      HSR_HIP(dbuf_reserve(c->s.host_in[a], sizeof(float) * (size_t)std::max<int64_t>(count, 1)));
      float *b = dptr<float>(c->s.host_in[a]);
      if (count) HSR_HIP(hipMemcpyAsync(b, d[a], sizeof(float) * count, hipMemcpyHostToDevice, st));
      d[a] = b;
    }
  }
  HSR_HIP(hipEventRecord(c->sub.ev[0], st));
  if ((s = drift_run(c, count, k, d[0], d[1], d[2], d[3], d[4], d[5]))) return s;
  HSR_HIP(hipEventRecord(c->sub.ev[1], st));
  if (host && count) {
    HSR_HIP(hipMemcpyAsync(x, d[0], sizeof(float) * count, hipMemcpyDeviceToHost, st));
    HSR_HIP(hipMemcpyAsync(y, d[1], sizeof(float) * count, hipMemcpyDeviceToHost, st));
    HSR_HIP(hipMemcpyAsync(z, d[2], sizeof(float) * count, hipMemcpyDeviceToHost, st));
  }
  HSR_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->sub.ev[0], c->sub.ev[1]));
  c->sub.drift_ms = ms;
  return HSR_OK;
}

hsr_status hsr_resort(hsr_ctx *c, int64_t count, const int32_t ng[3], float *x,
                      float *y, float *z, float *vx, float *vy, float *vz,
                      float *mass, float *phi, int64_t *id, uint16_t *mask,
                      int64_t *inside, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!x || !y || !z || !inside) {
    g_err = "hsr_resort: null positions or output";
    return HSR_ERR_ARG;
  }
  if ((s = check_count_grid("hsr_resort", count, ng))) return s;
  if ((s = sub_events(c))) return s;
  float *f[8] = {x, y, z, vx, vy, vz, mass, phi};
  float *df[8];
  int64_t *did = id;
  uint16_t *dmask = mask;
  const bool host = !(ptr_flags & HSR_DEVICE_PTRS);
  if (host) {
    if ((s = stage_in(c, c->s.host_in, count, f, id, mask, df, &did, &dmask))) return s;
  } else {
    for (int a = 0; a < 8; ++a) df[a] = f[a];
  }
  HSR_HIP(hipEventRecord(c->sub.ev[2], c->stream));
  if ((s = resort_run(c, count, ng, df, did, dmask, inside))) return s;
  HSR_HIP(hipEventRecord(c->sub.ev[3], c->stream));
  if (host && (s = stage_out(c, count, f, id, mask, df, did, dmask))) return s;
  HSR_HIP(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->sub.ev[2], c->sub.ev[3]));
  c->sub.resort_ms = ms;
  return HSR_OK;
}

hsr_status hsr_resort_noalloc(hsr_ctx *c, int64_t count, const int32_t ng[3], float *x,
                              float *y, float *z, float *vx, float *vy, float *vz,
                              float *mass, float *phi, int64_t *id, uint16_t *mask,
                              int64_t *inside, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!x || !y || !z || !inside) {
    g_err = "hsr_resort_noalloc: null positions or output";
    return HSR_ERR_ARG;
  }
  if (!(ptr_flags & HSR_DEVICE_PTRS)) {
    g_err = "hsr_resort_noalloc: device pointers only (HSR_DEVICE_PTRS); host arrays need "
            "staging, which hsr_resort allocates";
    return HSR_ERR_ARG;
  }
  if ((s = check_count_grid("hsr_resort_noalloc", count, ng))) return s;
  if (count > HSR_RESORT_NOALLOC_MAX) {
    g_err = "hsr_resort_noalloc: " + std::to_string(count) + " particles, above the scratch " +
            "array's limit of " + std::to_string(HSR_RESORT_NOALLOC_MAX) + " (use hsr_resort)";
    return HSR_ERR_ARG;
  }
  float *f[8] = {x, y, z, vx, vy, vz, mass, phi};
  if ((s = resort_noalloc_run(c, count, ng, f, id, mask, inside))) return s;
  return HSR_OK;
}

hsr_status hsr_cic(hsr_ctx *c, int64_t count, const float *x, const float *y,
                   const float *z, const int32_t ng[3], float gpscal, float *rho,
                   int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!x || !y || !z || !rho) {
    g_err = "hsr_cic: null positions or grid";
    return HSR_ERR_ARG;
  }
  if ((s = check_count_grid("hsr_cic", count, ng))) return s;
  if (count > (int64_t)(INT32_MAX / 8)) {
    g_err = "hsr_cic: at most 2^28 particles per call";
    return HSR_ERR_ARG;
  }
  if ((s = sub_events(c))) return s;
  hipStream_t st = c->stream;
  const int64_t ncell = (int64_t)ng[0] * ng[1] * ng[2] + 1;
  const float *d[3] = {x, y, z};
  float *drho = rho;
  const bool host = !(ptr_flags & HSR_DEVICE_PTRS);
  if (host) {
    const size_t nn = (size_t)std::max<int64_t>(count, 1);
    for (int a = 0; a < 3; ++a) {
      HSR_HIP(dbuf_reserve(c->s.host_in[a], sizeof(float) * nn));
      float *b = dptr<float>(c->s.host_in[a]);
      if (count) HSR_HIP(hipMemcpyAsync(b, d[a], sizeof(float) * count, hipMemcpyHostToDevice, st));
      d[a] = b;
    }
    HSR_HIP(dbuf_reserve(c->s.host_in[3], sizeof(float) * ncell));
    drho = dptr<float>(c->s.host_in[3]);
  }
  HSR_HIP(hipEventRecord(c->sub.ev[0], st));
  if ((s = cic_run(c, count, d[0], d[1], d[2], ng, gpscal, drho))) return s;
  HSR_HIP(hipEventRecord(c->sub.ev[1], st));
  if (host) HSR_HIP(hipMemcpyAsync(rho, drho, sizeof(float) * ncell, hipMemcpyDeviceToHost, st));
  HSR_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->sub.ev[0], c->sub.ev[1]));
  c->sub.cic_ms = ms;
  return HSR_OK;
}

hsr_status hsr_inverse_cic(hsr_ctx *c, int64_t count, const float *x, const float *y,
                           const float *z, const int32_t ng[3], const float *grid, float tau,
                           float fscal, float *v, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!x || !y || !z || !grid || !v) {
    g_err = "hsr_inverse_cic: null positions, grid or velocities";
    return HSR_ERR_ARG;
  }
  if ((s = check_count_grid("hsr_inverse_cic", count, ng))) return s;
  if ((s = sub_events(c))) return s;
  hipStream_t st = c->stream;
  const int64_t ncell = (int64_t)ng[0] * ng[1] * ng[2] + 1;
  const float *d[3] = {x, y, z};
  const float *dg = grid;
  float *dv = v;
  const bool host = !(ptr_flags & HSR_DEVICE_PTRS);
  if (host) {
    const size_t nn = (size_t)std::max<int64_t>(count, 1);
    for (int a = 0; a < 3; ++a) {
      HSR_HIP(dbuf_reserve(c->s.host_in[a], sizeof(float) * nn));
      float *b = dptr<float>(c->s.host_in[a]);
      if (count) HSR_HIP(hipMemcpyAsync(b, d[a], sizeof(float) * count, hipMemcpyHostToDevice, st));
      d[a] = b;
    }
    HSR_HIP(dbuf_reserve(c->s.host_in[3], sizeof(float) * ncell));
    float *bg = dptr<float>(c->s.host_in[3]);
    HSR_HIP(hipMemcpyAsync(bg, grid, sizeof(float) * ncell, hipMemcpyHostToDevice, st));
    dg = bg;
    HSR_HIP(dbuf_reserve(c->s.host_in[4], sizeof(float) * nn));
    dv = dptr<float>(c->s.host_in[4]);
    if (count) HSR_HIP(hipMemcpyAsync(dv, v, sizeof(float) * count, hipMemcpyHostToDevice, st));
  }
  HSR_HIP(hipEventRecord(c->sub.ev[2], st));
  if ((s = inverse_cic_run(c, count, d[0], d[1], d[2], ng, dg, tau, fscal, dv))) return s;
  HSR_HIP(hipEventRecord(c->sub.ev[3], st));
  if (host && count) HSR_HIP(hipMemcpyAsync(v, dv, sizeof(float) * count, hipMemcpyDeviceToHost, st));
  HSR_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->sub.ev[2], c->sub.ev[3]));
  c->sub.inverse_cic_ms = ms;
  return HSR_OK;
}

// Particles::map2, RCB monopole branch (Particles.cxx:1205-1280), on device
// arrays: resort all particles, set every mass to 1, build the tree over the
// m_Np_last particles inside the grid with the tree box [0, max ng), list the
// force box [edge, ng - edge) and kick with
// c = gpscal^3 / (4 pi) * fscal * tau * stepFraction.
static hsr_status map2_device(hsr_ctx *c, const hsr_subcycle_params *sp, double stepFraction,
                              int64_t count, float *const df[8], int64_t *did,
                              uint16_t *dmask, int64_t *np_last, double *resort_ms) {
  hsr_status s;
  hipStream_t st = c->stream;
  const float cc = map2_coeff(sp->gpscal, sp->fscal, sp->tau, stepFraction);
  c->sub.kick_coeff = cc;
  const int ngmax = std::max(std::max(sp->ng[0], sp->ng[1]), sp->ng[2]);
  const float tmin[3] = {0.f, 0.f, 0.f};
  const float tmax[3] = {(float)(1.0 * ngmax), (float)(1.0 * ngmax), (float)(1.0 * ngmax)};
  float flo[3], fhi[3];
  for (int d = 0; d < 3; ++d) {
    flo[d] = sp->edge;
    fhi[d] = (float)(1.0 * sp->ng[d] - sp->edge);
  }
  float *f[8];
  for (int a = 0; a < 8; ++a) f[a] = df[a];
  HSR_HIP(hipEventRecord(c->sub.ev[2], st));
  if ((s = resort_run(c, count, sp->ng, f, did, dmask, np_last))) return s;  // synchronizes
  HSR_HIP(hipEventRecord(c->sub.ev[3], st));
  if ((s = fill_run(c, count, 1.0f, f[6]))) return s;
  if ((s = hsr_build(c, *np_last, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], did, dmask,
                     tmin, tmax, HSR_DEVICE_PTRS)))
    return s;
  if ((s = hsr_lists(c, flo, fhi))) return s;
  if ((s = hsr_kick(c, cc))) return s;  // synchronizes
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->sub.ev[2], c->sub.ev[3]));
  *resort_ms += ms;
  return HSR_OK;
}

hsr_status hsr_map2(hsr_ctx *c, const hsr_subcycle_params *sp, double stepFraction,
                    int64_t count, float *x, float *y, float *z, float *vx, float *vy,
                    float *vz, float *mass, float *phi, int64_t *id, uint16_t *mask,
                    int64_t *inside, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!sp || !x || !y || !z || !vx || !vy || !vz || !mass || !inside) {
    g_err = "hsr_map2: null parameters, positions, velocities, masses or output";
    return HSR_ERR_ARG;
  }
  if ((s = check_count_grid("hsr_map2", count, sp->ng))) return s;
  if ((s = sub_events(c))) return s;
  float *f[8] = {x, y, z, vx, vy, vz, mass, phi};
  float *df[8];
  int64_t *did = id;
  uint16_t *dmask = mask;
  const bool host = !(ptr_flags & HSR_DEVICE_PTRS);
  if (host) {
    if ((s = stage_in(c, c->sub.arr, count, f, id, mask, df, &did, &dmask))) return s;
  } else {
    for (int a = 0; a < 8; ++a) df[a] = f[a];
  }
  double resort_ms = 0;
  if ((s = map2_device(c, sp, stepFraction, count, df, did, dmask, inside, &resort_ms))) return s;
  if (host && (s = stage_out(c, count, f, id, mask, df, did, dmask))) return s;
  HSR_HIP(hipStreamSynchronize(c->stream));
  c->sub.resort_ms = resort_ms;
  return HSR_OK;
}

hsr_status hsr_subcycle(hsr_ctx *c, const hsr_subcycle_params *sp, int64_t count,
                        float *x, float *y, float *z, float *vx, float *vy,
                        float *vz, float *mass, float *phi, int64_t *id,
                        uint16_t *mask, int64_t *inside, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!sp || !x || !y || !z || !vx || !vy || !vz || !mass || !inside) {
    g_err = "hsr_subcycle: null parameters, positions, velocities, masses or output";
    return HSR_ERR_ARG;
  }
  if (sp->nsub < 1) {
    g_err = "hsr_subcycle: nsub must be >= 1";
    return HSR_ERR_ARG;
  }
  if ((s = check_count_grid("hsr_subcycle", count, sp->ng))) return s;
  if ((s = sub_events(c))) return s;
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = c->stream;
  float *f[8] = {x, y, z, vx, vy, vz, mass, phi};
  float *df[8];
  int64_t *did = id;
  uint16_t *dmask = mask;
  const bool host = !(ptr_flags & HSR_DEVICE_PTRS);
  if (host) {
    if ((s = stage_in(c, c->sub.arr, count, f, id, mask, df, &did, &dmask))) return s;
  } else {
    for (int a = 0; a < 8; ++a) df[a] = f[a];
  }
  // Particles::subCycle (Particles.cxx:1176-1201): the TimeStepper is not
  // advanced inside the cycle, so every sub-step uses the same factors;
  // map1 receives pp, stepFraction * tau2 and adot as floats.
  const double stepFraction = 1.0 / sp->nsub;
  const float k = map1_factor((float)sp->pp, (float)(stepFraction * sp->tau2), (float)sp->adot,
                              sp->alpha);
  c->sub.drift_factor = k;
  double drift_ms = 0, resort_ms = 0;
  int64_t np_last = -1;
  auto timed = [&](int e0, int e1, double &acc) -> hsr_status {
    float ms = 0.f;
    HSR_HIP(hipEventElapsedTime(&ms, c->sub.ev[e0], c->sub.ev[e1]));
    acc += ms;
    return HSR_OK;
  };
  for (int step = 0; step < sp->nsub; ++step) {
    if (!sp->skip_stream) {
      HSR_HIP(hipEventRecord(c->sub.ev[0], st));
      if ((s = drift_run(c, count, k, df[0], df[1], df[2], df[3], df[4], df[5]))) return s;
      HSR_HIP(hipEventRecord(c->sub.ev[1], st));
    }
    if (!sp->skip_kick &&
        (s = map2_device(c, sp, stepFraction, count, df, did, dmask, &np_last, &resort_ms)))
      return s;
    if (!sp->skip_stream) {
      HSR_HIP(hipEventRecord(c->sub.ev[4], st));
      if ((s = drift_run(c, count, k, df[0], df[1], df[2], df[3], df[4], df[5]))) return s;
      HSR_HIP(hipEventRecord(c->sub.ev[5], st));
    }
    HSR_HIP(hipStreamSynchronize(st));
    if (!sp->skip_stream) {
      if ((s = timed(0, 1, drift_ms)) || (s = timed(4, 5, drift_ms))) return s;
    }
  }
  if (host && (s = stage_out(c, count, f, id, mask, df, did, dmask))) return s;
  HSR_HIP(hipStreamSynchronize(st));
  *inside = np_last;
  c->sub.drift_ms = drift_ms;
  c->sub.resort_ms = resort_ms;
  c->sub.subcycle_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return HSR_OK;
}

}  // extern "C"

static hsr_status check_domain(const char *who, const hsr_domain *d) {
  if (!d || !(d->rL > 0.f) || !std::isfinite(d->rL) || d->ng < 1 || !(d->oL_in >= 0.f) ||
      !(d->oL_in <= d->rL)) {
    g_err = std::string(who) + ": domain needs rL > 0, ng >= 1 and 0 <= oL <= rL";
    return HSR_ERR_ARG;
  }
  return HSR_OK;
}

hsr_status hsr_domain_init(hsr_domain *d) {
  hsr_status s = check_domain("hsr_domain_init", d);
  if (s) return s;
  domain_derive(d);
  const int64_t ngt = (int64_t)d->ng + 2 * (int64_t)d->ng_overload + 1;
  if (d->ng_overload < 0 || ngt > (1 << 24) || ngt * ngt * ngt >= (int64_t)INT32_MAX) {
    g_err = "hsr_domain_init: ng + 2 ng_overload + 1 must stay below 1291 (cells < 2^31)";
    return HSR_ERR_ARG;
  }
  return HSR_OK;
}

hsr_status hsr_refresh(hsr_ctx *c, const hsr_domain *dom_in, float anow, int64_t *count,
                       int64_t capacity, float *x, float *y, float *z, float *vx, float *vy,
                       float *vz, float *mass, float *phi, int64_t *id, uint16_t *mask,
                       int64_t *inside, int ptr_flags) {
  hsr_status s = check_ctx(c);
  if (s) return s;
  if (!count || !inside || !x || !y || !z || !vx || !vy || !vz || !mass || !phi || !id ||
      !mask) {
    g_err = "hsr_refresh: every particle array, count and inside are required";
    return HSR_ERR_ARG;
  }
  hsr_domain dom = dom_in ? *dom_in : hsr_domain{};
  if ((s = hsr_domain_init(&dom))) return s;
  if (!(anow > 0.f) || !std::isfinite(anow)) {
    g_err = "hsr_refresh: anow must be positive and finite";
    return HSR_ERR_ARG;
  }
  const int64_t n = *count;
  if (n < 0 || n >= (int64_t)INT32_MAX || capacity < n) {
    g_err = "hsr_refresh: need 0 <= count <= capacity and count < 2^31";
    return HSR_ERR_ARG;
  }
  if ((s = sub_events(c))) return s;
  hipStream_t st = c->stream;
  float *f[8] = {x, y, z, vx, vy, vz, mass, phi};
  PArrays in{};
  const bool host = !(ptr_flags & HSR_DEVICE_PTRS);
  if (host) {
    if ((s = stage_in(c, c->s.host_in, n, f, id, mask, in.f, &in.id, &in.mask))) return s;
  } else {
    for (int a = 0; a < 8; ++a) in.f[a] = f[a];
    in.id = id;
    in.mask = mask;
  }
  HSR_HIP(hipEventRecord(c->sub.ev[4], st));
  int64_t n_out = 0, alive = 0;
  if ((s = refresh_count(c, &dom, anow, n, in.f[0], in.f[1], in.f[2], &n_out, &alive))) return s;
  if (n_out > capacity || n_out >= (int64_t)INT32_MAX) {
    g_err = "hsr_refresh: " + std::to_string(n_out) + " particles after the exchange exceed the capacity " +
            std::to_string(capacity) + " (or 2^31)";
    *count = n_out;
    return HSR_ERR_ARG;
  }
  PArrays out{};
  if (host) {
    const size_t no = (size_t)std::max<int64_t>(n_out, 1);
    HSR_HIP(dbuf_reserve(c->sub.xout, (8 + 8 * 4 + 2) * no + 256));
    char *p = (char *)c->sub.xout.p;
    out.id = (int64_t *)p;
    p += sizeof(int64_t) * no;
    for (int a = 0; a < 8; ++a) {
      out.f[a] = (float *)p;
      p += sizeof(float) * no;
    }
    out.mask = (uint16_t *)p;
  } else {
    out = in;
  }
  int64_t ins = 0;
  if ((s = refresh_finish(c, &dom, anow, n, in, n_out, out, &ins))) return s;
  HSR_HIP(hipEventRecord(c->sub.ev[5], st));
  if (host && (s = stage_out(c, n_out, f, id, mask, out.f, out.id, out.mask))) return s;
  HSR_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  HSR_HIP(hipEventElapsedTime(&ms, c->sub.ev[4], c->sub.ev[5]));
  c->sub.refresh_ms = ms;
  c->sub.refresh_alive = alive;
  c->sub.refresh_dead = n_out - alive;
  *count = n_out;
  *inside = ins;
  return HSR_OK;
}
