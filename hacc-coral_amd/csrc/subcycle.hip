// subcycle.hip -- the particle steps either side of the short-range kick in
// the sub-cycle (Particles::subCycle, Particles.cxx:1176-1201), on the device
// so particles stay resident across the sub-steps.
//
//   drift   x += k v with k = prefactor * tau rounded once on the host, then
//           (k * v) rounded and added (Particles::map1, Particles.cxx:732-758;
//           built without FMA contraction so every element rounds as there).
//   resort  stable counting sort of the particles by unit cell
//           (Particles::resortParticles, Particles.cxx:402-488): cell
//           (floor x * ng1 + floor y) * ng2 + floor z inside the ng grid,
//           every other particle into one tail bucket after the last cell,
//           input order kept inside a cell.  The ten particle arrays are
//           permuted in place; the count of particles inside the grid is
//           the reference's m_Np_last.
//
// The sort is a stable LSD radix sort of (cell, index) pairs over the bits
// the cell count needs (same permutation as the reference's counting pass:
// a destination is the number of particles of lower cells plus the number of
// earlier particles of the same cell), followed by one gather of all ten
// arrays and a copy back.  Every kernel here is a streaming pass bounded by
// HBM bandwidth.
#include "hsr_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>
#include <string>

namespace {

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__global__ void __launch_bounds__(256)
k_drift4(int64_t n4, float k, float4 *__restrict__ x, float4 *__restrict__ y,
         float4 *__restrict__ z, const float4 *__restrict__ vx,
         const float4 *__restrict__ vy, const float4 *__restrict__ vz) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n4) return;
  float4 a = x[i], v = vx[i];
  a.x = a.x + k * v.x; a.y = a.y + k * v.y; a.z = a.z + k * v.z; a.w = a.w + k * v.w;
  x[i] = a;
  a = y[i]; v = vy[i];
  a.x = a.x + k * v.x; a.y = a.y + k * v.y; a.z = a.z + k * v.z; a.w = a.w + k * v.w;
  y[i] = a;
  a = z[i]; v = vz[i];
  a.x = a.x + k * v.x; a.y = a.y + k * v.y; a.z = a.z + k * v.z; a.w = a.w + k * v.w;
  z[i] = a;
}

__global__ void __launch_bounds__(256)
k_drift1(int64_t i0, int64_t n, float k, float *__restrict__ x, float *__restrict__ y,
         float *__restrict__ z, const float *__restrict__ vx,
         const float *__restrict__ vy, const float *__restrict__ vz) {
  const int64_t i = i0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = x[i] + k * vx[i];
  y[i] = y[i] + k * vy[i];
  z[i] = z[i] + k * vz[i];
}

// Cell key of every particle.  floorf and float comparisons against the
// (exactly representable) grid sizes decide membership: for finite values
// this equals the reference's int test on (int)floorf(x), and NaN or values
// beyond the int range land in the tail bucket as the x86 conversion
// (INT_MIN) sends them there.
__global__ void __launch_bounds__(256)
k_cell_keys(int64_t n, const float *__restrict__ x, const float *__restrict__ y,
            const float *__restrict__ z, int ng0, int ng1, int ng2, uint32_t tail,
            uint32_t *__restrict__ key, int32_t *__restrict__ val,
            unsigned long long *__restrict__ outside) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool out = false;
  if (i < n) {
    const float fx = floorf(x[i]), fy = floorf(y[i]), fz = floorf(z[i]);
    const bool in = fx >= 0.f && fx < (float)ng0 && fy >= 0.f && fy < (float)ng1 &&
                    fz >= 0.f && fz < (float)ng2;
    key[i] = in ? ((uint32_t)fx * (uint32_t)ng1 + (uint32_t)fy) * (uint32_t)ng2 + (uint32_t)fz
                : tail;
    val[i] = (int32_t)i;
    out = !in;
  }
  const unsigned long long b = __ballot(out);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(outside, (unsigned long long)__popcll(b));
}

struct Ten {
  float *f[8];
  int64_t *id;
  uint16_t *mask;
};

// out[k] = in[src[k]] for each present array.
__global__ void __launch_bounds__(256)
k_gather_ten(int64_t n, const int32_t *__restrict__ src, Ten in, Ten out) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t s = src[k];
#pragma unroll
  for (int a = 0; a < 8; ++a)
    if (in.f[a]) out.f[a][k] = in.f[a][s];
  if (in.id) out.id[k] = in.id[s];
  if (in.mask) out.mask[k] = in.mask[s];
}

__global__ void k_fill_f32(int64_t n, float v, float *a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}

__global__ void k_zero_ull(unsigned long long *p) {
  if (threadIdx.x == 0) p[0] = 0ull;
}

}  // namespace

hsr_status drift_run(hsr_ctx *c, int64_t n, float k, float *x, float *y, float *z,
                     const float *vx, const float *vy, const float *vz) {
  if (n <= 0) return HSR_OK;
  hipStream_t st = c->stream;
  const void *p[6] = {x, y, z, vx, vy, vz};
  bool al = true;
  for (const void *q : p) al = al && ((uintptr_t)q % 16 == 0);
  int64_t n4 = al ? n / 4 : 0;
  if (n4)
    k_drift4<<<nblk(n4, 256), 256, 0, st>>>(n4, k, (float4 *)x, (float4 *)y, (float4 *)z,
                                          (const float4 *)vx, (const float4 *)vy,
                                          (const float4 *)vz);
  const int64_t i0 = 4 * n4;
  if (i0 < n)
    k_drift1<<<nblk(n - i0, 256), 256, 0, st>>>(i0, n, k, x, y, z, vx, vy, vz);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}

static int cell_key_bits(const int32_t ng[3]) {
  const uint64_t Ng = (uint64_t)ng[0] * (uint64_t)ng[1] * (uint64_t)ng[2];
  int bits = 1;
  while (bits < 32 && (Ng >> bits) != 0) ++bits;  // keys 0..Ng need this many bits
  return bits;
}

// Bytes of hipcub radix-sort temporary storage for n (cell, index) pairs.
static hsr_status sort_tmp_bytes(int64_t n, int bits, hipStream_t st, size_t *tmp) {
  *tmp = 0;
  uint32_t *k = nullptr;
  int32_t *v = nullptr;
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, *tmp, k, k, v, v, (int)n, 0, bits, st));
  return HSR_OK;
}

// Cell keys of n particles and their stable radix sort, in the caller's
// scratch: `keys` holds 4n words (key, key', val, val') followed by the
// 8-byte outside counter; `tmp` holds tmp_bytes of sort storage.  Leaves the
// source index of each destination in *src and the number of particles
// outside the grid in *outside_dev (device).
static hsr_status sort_cells(hsr_ctx *c, int64_t n, const int32_t ng[3], const float *x,
                             const float *y, const float *z, uint32_t *keys, void *tmp,
                             size_t tmp_bytes, const int32_t **src,
                             unsigned long long **outside_dev) {
  hipStream_t st = c->stream;
  const uint64_t Ng = (uint64_t)ng[0] * (uint64_t)ng[1] * (uint64_t)ng[2];
  const int bits = cell_key_bits(ng);
  const size_t nn = (size_t)n;
  uint32_t *key = keys;
  uint32_t *key2 = key + nn;
  int32_t *val = (int32_t *)(key2 + nn);
  int32_t *val2 = val + nn;
  unsigned long long *outside = (unsigned long long *)(val2 + nn);  // 16n bytes in: aligned
  k_zero_ull<<<1, 64, 0, st>>>(outside);
  k_cell_keys<<<nblk(n, 256), 256, 0, st>>>(n, x, y, z, ng[0], ng[1], ng[2], (uint32_t)Ng,
                                          key, val, outside);
  HSR_HIP(hipGetLastError());
  size_t t = tmp_bytes;
  HSR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, t, key, key2, val, val2, (int)n, 0, bits, st));
  *src = val2;
  *outside_dev = outside;
  return HSR_OK;
}

// sort_cells over the context's buffers (grown as needed).
static hsr_status sort_cells_ctx(hsr_ctx *c, int64_t n, const int32_t ng[3], const float *x,
                                 const float *y, const float *z, const int32_t **src,
                                 unsigned long long **outside_dev) {
  size_t tmp = 0;
  hsr_status s = sort_tmp_bytes(n, cell_key_bits(ng), c->stream, &tmp);
  if (s) return s;
  // key, key', val, val' (4 x 4 B) and the outside counter
  HSR_HIP(dbuf_reserve(c->sub.keys, 16 * (size_t)n + 256));
  HSR_HIP(dbuf_reserve(c->s.cub, tmp));
  return sort_cells(c, n, ng, x, y, z, dptr<uint32_t>(c->sub.keys), c->s.cub.p, tmp, src,
                    outside_dev);
}

hsr_status resort_into(hsr_ctx *c, int64_t n, const int32_t ng[3], const PArrays &in,
                       const PArrays &out, int64_t *inside) {
  hipStream_t st = c->stream;
  *inside = 0;
  if (n <= 0) return HSR_OK;
  const int32_t *src = nullptr;
  unsigned long long *outside = nullptr;
  hsr_status s = sort_cells_ctx(c, n, ng, in.f[0], in.f[1], in.f[2], &src, &outside);
  if (s) return s;
  Ten ti{}, to{};
  for (int a = 0; a < 8; ++a) {
    ti.f[a] = in.f[a];
    to.f[a] = out.f[a];
  }
  ti.id = in.id;
  to.id = out.id;
  ti.mask = in.mask;
  to.mask = out.mask;
  k_gather_ten<<<nblk(n, 256), 256, 0, st>>>(n, src, ti, to);
  HSR_HIP(hipGetLastError());
  unsigned long long h_out = 0;
  HSR_HIP(hipMemcpyAsync(&h_out, outside, sizeof(h_out), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipStreamSynchronize(st));
  *inside = n - (int64_t)h_out;
  return HSR_OK;
}

static size_t stage_bytes(int64_t n, float *const f[8], const int64_t *id, const uint16_t *mask) {
  const size_t nn = (size_t)n;
  size_t bytes = 0;
  for (int a = 0; a < 8; ++a)
    if (f[a]) bytes += sizeof(float) * nn;
  if (id) bytes += sizeof(int64_t) * nn;
  if (mask) bytes += sizeof(uint16_t) * nn;
  return bytes;
}

// Gather the present arrays into `stage` in the sorted order, copy them back
// in place and read the outside count.
static hsr_status gather_back(hsr_ctx *c, int64_t n, const int32_t *src,
                              const unsigned long long *outside, float *const f[8], int64_t *id,
                              uint16_t *mask, char *sp, int64_t *inside) {
  hipStream_t st = c->stream;
  const size_t nn = (size_t)n;
  Ten in{}, out{};
  if (id) {  // 8-byte arrays first keeps every staging array aligned
    in.id = id;
    out.id = (int64_t *)sp;
    sp += sizeof(int64_t) * nn;
  }
  for (int a = 0; a < 8; ++a)
    if (f[a]) {
      in.f[a] = f[a];
      out.f[a] = (float *)sp;
      sp += sizeof(float) * nn;
    }
  if (mask) {
    in.mask = mask;
    out.mask = (uint16_t *)sp;
  }
  k_gather_ten<<<nblk(n, 256), 256, 0, st>>>(n, src, in, out);
  HSR_HIP(hipGetLastError());
  for (int a = 0; a < 8; ++a)
    if (f[a]) HSR_HIP(hipMemcpyAsync(f[a], out.f[a], sizeof(float) * nn, hipMemcpyDeviceToDevice, st));
  if (id) HSR_HIP(hipMemcpyAsync(id, out.id, sizeof(int64_t) * nn, hipMemcpyDeviceToDevice, st));
  if (mask) HSR_HIP(hipMemcpyAsync(mask, out.mask, sizeof(uint16_t) * nn, hipMemcpyDeviceToDevice, st));
  unsigned long long h_out = 0;
  HSR_HIP(hipMemcpyAsync(&h_out, outside, sizeof(h_out), hipMemcpyDeviceToHost, st));
  HSR_HIP(hipStreamSynchronize(st));
  *inside = n - (int64_t)h_out;
  return HSR_OK;
}

hsr_status resort_run(hsr_ctx *c, int64_t n, const int32_t ng[3], float *f[8],
                      int64_t *id, uint16_t *mask, int64_t *inside) {
  *inside = 0;
  if (n <= 0) return HSR_OK;
  const int32_t *val2 = nullptr;
  unsigned long long *outside = nullptr;
  hsr_status s = sort_cells_ctx(c, n, ng, f[0], f[1], f[2], &val2, &outside);
  if (s) return s;
  HSR_HIP(dbuf_reserve(c->sub.stage, stage_bytes(n, f, id, mask) + 256));
  return gather_back(c, n, val2, outside, f, id, mask, (char *)c->sub.stage.p, inside);
}

// ---------------------------------------------------------------------------
// Allocation-free resort (hsr_resort_noalloc).  Its scratch is one
// file-scope device array, `scratch`, of fixed size, held by every device the
// library runs on until the process ends:
//   [keys: 16 B per particle + 256 B][staging: 42 B per particle + 256 B]
//   [radix-sort storage: the rest]
// for at most HSR_RESORT_NOALLOC_MAX particles.  Host code never passes the
// array's name to a kernel or copy (on the host it is only a stand-in
// symbol): each device's address comes from hipGetSymbolAddress, taken once
// per device with that device current, and is kept in g_scratch_dev.
namespace {
constexpr size_t kNoallocKeys = 16ull * HSR_RESORT_NOALLOC_MAX + 256;
constexpr size_t kNoallocStage = 42ull * HSR_RESORT_NOALLOC_MAX + 256;
constexpr size_t kNoallocSort = 64ull << 20;
constexpr size_t kNoallocBytes = kNoallocKeys + kNoallocStage + kNoallocSort;
static_assert(kNoallocBytes == HSR_RESORT_NOALLOC_SCRATCH_BYTES, "scratch size");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_dev[kMaxDevices];
}  // namespace

__device__ __align__(256) unsigned char scratch[kNoallocBytes];

hsr_status noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    hsr_set_error("scratch set-up: device ordinal beyond the library's table");
    return HSR_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return HSR_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    hsr_set_error(std::string("scratch set-up: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    return HSR_ERR_HIP;
  }
  g_scratch_dev[device] = (char *)p;
  return HSR_OK;
}

hsr_status resort_noalloc_run(hsr_ctx *c, int64_t n, const int32_t ng[3], float *f[8],
                              int64_t *id, uint16_t *mask, int64_t *inside) {
  *inside = 0;
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_dev[c->device] : nullptr;
  }
  if (!base) {
    hsr_set_error("hsr_resort_noalloc: the device's scratch address was not set up");
    return HSR_ERR_STATE;
  }
  if (n <= 0) return HSR_OK;
  size_t tmp = 0;
  hsr_status s = sort_tmp_bytes(n, cell_key_bits(ng), c->stream, &tmp);
  if (s) return s;
  if (tmp > kNoallocSort) {
    hsr_set_error("hsr_resort_noalloc: the radix sort needs more storage than the scratch array holds");
    return HSR_ERR_UNSUPPORTED;
  }
  uint32_t *keys = (uint32_t *)base;
  char *stage = base + kNoallocKeys;
  void *sort_tmp = base + kNoallocKeys + kNoallocStage;
  const int32_t *val2 = nullptr;
  unsigned long long *outside = nullptr;
  if ((s = sort_cells(c, n, ng, f[0], f[1], f[2], keys, sort_tmp, tmp, &val2, &outside)))
    return s;
  return gather_back(c, n, val2, outside, f, id, mask, stage, inside);
}

hsr_status fill_run(hsr_ctx *c, int64_t n, float v, float *a) {
  if (n > 0) k_fill_f32<<<nblk(n, 256), 256, 0, c->stream>>>(n, v, a);
  HSR_HIP(hipGetLastError());
  return HSR_OK;
}
