// Particle tracer, particle quantizer, slice extractor and surface-pressure
// kernels, plus the library's error state and device query.
//
// Each kernel restates one reference kernel operation by operation; with the
// library compiled -ffp-contract=off, the integer outputs (cell indices, kill
// decisions, quantized bytes) are identical to a CPU evaluation of the same
// IEEE operations on the same inputs.
#include <hip/hip_runtime.h>

#include <hipcub/device/device_radix_sort.hpp>

#include <cstring>

#include "wt_device.h"
#include "wt_host.h"
#include "wt_particles.h"

namespace wt {

char* error_buffer() {
  static thread_local char buf[512] = {0};
  return buf;
}

constexpr int ABLOCK = 256;

// lbm.py:233-263
__global__ __launch_bounds__(ABLOCK) void k_advect(AdvectArgs a) {
  const uint32_t j = blockIdx.x * ABLOCK + threadIdx.x;
  if (j < a.num_p) advect_particle(a, j);
}

// Slot storage: out[j] = in[order[j]] for the particle arrays and their ids.
__global__ __launch_bounds__(ABLOCK) void k_gather(const float4* __restrict__ pos, const float* __restrict__ spd,
                                                   const uint32_t* __restrict__ ids,
                                                   const uint32_t* __restrict__ order, uint32_t num_p,
                                                   float4* __restrict__ pos_out, float* __restrict__ spd_out,
                                                   uint32_t* __restrict__ ids_out) {
  const uint32_t j = blockIdx.x * ABLOCK + threadIdx.x;
  if (j >= num_p) return;
  const uint32_t k = order[j];
  pos_out[j] = pos[k];
  spd_out[j] = spd[k];
  ids_out[j] = ids ? ids[k] : k;
}

// Sort key of each particle: the index of the cell its position falls in
// (clamped), so sorted order walks the tunnel in memory order.
__global__ __launch_bounds__(ABLOCK) void k_particle_keys(const float4* __restrict__ pos, uint32_t* keys,
                                                          uint32_t* ids, uint32_t num_p, int nx, int ny, int nz) {
  const uint32_t i = blockIdx.x * ABLOCK + threadIdx.x;
  if (i >= num_p) return;
  const float4 p = pos[i];
  const float qx = clampf(p.x, 0.0f, (float)nx - 1.0f), qy = clampf(p.y, 0.0f, (float)ny - 1.0f);
  const float qz = clampf(p.z, 0.0f, (float)nz - 1.0f);
  keys[i] = cidx((int)qx, (int)qy, (int)qz, nx, ny, nz);   // NaN -> 0 after the clamps
  ids[i] = i;
}

// lbm.py:266-275
__global__ __launch_bounds__(ABLOCK) void k_pack(const float4* __restrict__ pos, const float* __restrict__ spd,
                                                 const uint32_t* __restrict__ ids, ushort4* q16, uint8_t* q8,
                                                 uint32_t num_p, int nx, int ny, int nz) {
  const uint32_t j = blockIdx.x * ABLOCK + threadIdx.x;
  if (j >= num_p) return;
  // outputs are in particle-id order whatever slot holds the particle
  const uint32_t i = ids ? ids[j] : j;
  const float4 p = pos[j];
  ushort4 q;
  q.x = (uint16_t)(int)(clampf(p.x / (float)nx, 0.0f, 1.0f) * 65535.0f + 0.5f);
  q.y = (uint16_t)(int)(clampf(p.y / (float)ny, 0.0f, 1.0f) * 65535.0f + 0.5f);
  q.z = (uint16_t)(int)(clampf(p.z / (float)nz, 0.0f, 1.0f) * 65535.0f + 0.5f);
  q.w = (uint16_t)(int)clampf(p.w, 0.0f, 65535.0f);
  q16[i] = q;
  q8[i] = (uint8_t)(int)(clampf(spd[j] / 2.0f, 0.0f, 1.0f) * 255.0f + 0.5f);
}

struct SliceArgs {
  MacView mac;
  const uint8_t* cls;
  uint8_t* out;
  int nx, ny, nz, axis, index, field, w;
  uint32_t npix;
  float uin, lref, vmin, vmax;
};

// lbm.py:278-318
__global__ __launch_bounds__(ABLOCK) void k_slice(SliceArgs a) {
  const uint32_t k = blockIdx.x * ABLOCK + threadIdx.x;
  if (k >= a.npix) return;
  const int pa = (int)(k % (uint32_t)a.w), pb = (int)(k / (uint32_t)a.w);
  int cx = a.index, cy = pb, cz = pa;
  if (a.axis == 2) { cx = pa; cy = pb; cz = a.index; }
  else if (a.axis == 1) { cx = pa; cy = a.index; cz = pb; }
  const uint32_t idx = (uint32_t)cx + (uint32_t)a.nx * ((uint32_t)cy + (uint32_t)a.ny * (uint32_t)cz);
  const float4 m = a.mac[idx];
  float v = m.x / a.uin;
  if (a.field == 0) {
    v = sqrtf(m.x * m.x + m.y * m.y + m.z * m.z) / a.uin;
  } else if (a.field == 1) {
    v = (m.w - 1.0f) / (1.5f * a.uin * a.uin);
  } else if (a.field == 2) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const float4 xp = a.mac[cidx(cx + 1, cy, cz, nx, ny, nz)], xm = a.mac[cidx(cx - 1, cy, cz, nx, ny, nz)];
    const float4 yp = a.mac[cidx(cx, cy + 1, cz, nx, ny, nz)], ym = a.mac[cidx(cx, cy - 1, cz, nx, ny, nz)];
    const float4 zp = a.mac[cidx(cx, cy, cz + 1, nx, ny, nz)], zm = a.mac[cidx(cx, cy, cz - 1, nx, ny, nz)];
    const float dxy = (xp.y - xm.y) * 0.5f, dxz = (xp.z - xm.z) * 0.5f;
    const float dyx = (yp.x - ym.x) * 0.5f, dyz = (yp.z - ym.z) * 0.5f;
    const float dzx = (zp.x - zm.x) * 0.5f, dzy = (zp.y - zm.y) * 0.5f;
    const float wx = dyz - dzy, wy = dzx - dxz, wz = dxy - dyx;
    v = sqrtf(wx * wx + wy * wy + wz * wz) * a.lref / a.uin;
  }
  const float t = clampf((v - a.vmin) / (a.vmax - a.vmin), 0.0f, 1.0f);
  uchar2 o;
  o.x = (uint8_t)(int)(t * 255.0f + 0.5f);
  o.y = (a.cls[idx] & CLS_CODE) != 0 ? 0 : 255;
  reinterpret_cast<uchar2*>(a.out)[k] = o;
}

// Fluid-only trilinear density, lbm.py:321-350.
__device__ float2 rho_fluid(const MacView& mac, const uint8_t* __restrict__ cls, float px,
                            float py, float pz, int nx, int ny, int nz) {
  const float qx = clampf(px, 0.0f, (float)nx - 1.001f);
  const float qy = clampf(py, 0.0f, (float)ny - 1.001f);
  const float qz = clampf(pz, 0.0f, (float)nz - 1.001f);
  const int ix = (int)floorf(qx), iy = (int)floorf(qy), iz = (int)floorf(qz);
  const float tx = qx - (float)ix, ty = qy - (float)iy, tz = qz - (float)iz;
  float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
    const uint32_t ci = cidx(ix + ox, iy + oy, iz + oz, nx, ny, nz);
    const float wx = ox ? tx : 1.0f - tx;
    const float wy = oy ? ty : 1.0f - ty;
    const float wz = oz ? tz : 1.0f - tz;
    const float wt = wx * wy * wz;
    if ((cls[ci] & CLS_CODE) == 0) {
      s0 = s0 + wt * mac[ci].w;
      s1 = s1 + wt;
    }
  }
  return make_float2(s0, s1);
}

struct CpArgs {
  MacView mac;
  const uint8_t* cls;
  const float* verts;
  const float* norms;
  int8_t* out;
  uint32_t nv;
  int nx, ny, nz;
  float uin, cp_range;
};

// lbm.py:353-368
__global__ __launch_bounds__(ABLOCK) void k_cp(CpArgs a) {
  const uint32_t i = blockIdx.x * ABLOCK + threadIdx.x;
  if (i >= a.nv) return;
  const float px = a.verts[3 * (size_t)i], py = a.verts[3 * (size_t)i + 1], pz = a.verts[3 * (size_t)i + 2];
  const float rx = a.norms[3 * (size_t)i], ry = a.norms[3 * (size_t)i + 1], rz = a.norms[3 * (size_t)i + 2];
  const float2 A = rho_fluid(a.mac, a.cls, px + rx * 1.25f, py + ry * 1.25f, pz + rz * 1.25f, a.nx, a.ny, a.nz);
  const float2 B = rho_fluid(a.mac, a.cls, px - rx * 1.25f, py - ry * 1.25f, pz - rz * 1.25f, a.nx, a.ny, a.nz);
  float2 s = B;
  if (A.y >= B.y) s = A;
  float cp = 0.0f;
  if (s.y > 0.05f) cp = (s.x / s.y - 1.0f) / (1.5f * a.uin * a.uin);
  a.out[i] = (int8_t)(int)roundf(clampf(cp / a.cp_range, -1.0f, 1.0f) * 127.0f);
}

// Plain streaming copy, 16 B per lane, one element per thread (the fastest
// copy form measured on MI355X): the reference point for achievable HBM
// bandwidth and for scaling the hardware byte counters.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(ABLOCK) void k_stream_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                        uint64_t n16) {
  const uint64_t i = (uint64_t)blockIdx.x * ABLOCK + threadIdx.x;
  if (i < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Self-test of the BF16 store rounding: f2bf (the hardware conversion the step
// kernels use) against round-to-nearest-even in bit arithmetic (x + 0x7fff +
// lsb, NaN quieted as (x >> 16) | 0x40: the oracle's rounding), over every
// float bit pattern. counts[0]: non-NaN inputs that differ, counts[1]: NaN inputs.
__device__ __forceinline__ uint16_t f2bf_bits(uint32_t x) {
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}
__global__ __launch_bounds__(ABLOCK) void k_zero_counts(unsigned long long* counts) {
  if (threadIdx.x < 2) counts[threadIdx.x] = 0ull;
}
__global__ __launch_bounds__(ABLOCK) void k_selftest_bf16(unsigned long long* counts) {
  const uint64_t stride = (uint64_t)gridDim.x * ABLOCK;
  unsigned long long bad[2] = {0ull, 0ull};
  for (uint64_t v = (uint64_t)blockIdx.x * ABLOCK + threadIdx.x; v < (1ull << 32); v += stride) {
    const uint32_t x = (uint32_t)v;
    if (f2bf(__uint_as_float(x)) != f2bf_bits(x)) bad[(x & 0x7fffffffu) > 0x7f800000u ? 1 : 0]++;
  }
  if (bad[0]) atomicAdd(&counts[0], bad[0]);
  if (bad[1]) atomicAdd(&counts[1], bad[1]);
}

int check_small_grid(const wt_grid* g) {
  if (!g || !g->mac || !g->cls) return fail(-1, "null grid, mac or class buffer");
  if (g->nx < 3 || g->ny < 3 || g->nz < 3) return fail(-1, "grid dimensions must be at least 3");
  if ((int64_t)g->nx * g->ny * g->nz >= ((int64_t)1 << 31)) return fail(-1, "grid too large");
  return 0;
}

}  // namespace wt

using namespace wt;

extern "C" {

const char* wt_last_error(void) { return error_buffer(); }

int wt_version(void) { return 1; }

int wt_device_info(char* name, int name_len, int64_t* total_bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(-2, "hipGetDevice failed");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(-2, "hipGetDeviceProperties failed");
  if (name && name_len > 0) {
    strncpy(name, prop.name, (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (total_bytes) *total_bytes = (int64_t)prop.totalGlobalMem;
  return 0;
}

int wt_stream_copy(void* dst, const void* src, int64_t bytes, int blocks, void* stream) {
  (void)blocks;
  if (!dst || !src || bytes <= 0 || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return fail(-1, "stream copy needs non-null 16-byte aligned buffers and a positive multiple of 16 bytes");
  const uint64_t n16 = (uint64_t)bytes / 16;
  if ((n16 + ABLOCK - 1) / ABLOCK >= (1ull << 31)) return fail(-1, "stream copy too large");
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)((n16 + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0,
                     (hipStream_t)stream, (const u32x4*)src, (u32x4*)dst, n16);
  return check_launch("wt_stream_copy");
}

int wt_selftest_bf16(uint64_t* counts_dev, void* stream) {
  if (!counts_dev || ((uintptr_t)counts_dev & 7)) return fail(-1, "counts need a non-null 8-byte aligned device buffer");
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* c = reinterpret_cast<unsigned long long*>(counts_dev);
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(ABLOCK), 0, s, c);
  hipLaunchKernelGGL(k_selftest_bf16, dim3(16384), dim3(ABLOCK), 0, s, c);
  return check_launch("wt_selftest_bf16");
}

int wt_advect(const wt_grid* g, float* pos, float* spd, int64_t num_p, float dt, float uin, float max_age,
              uint32_t seed, const wt_seed_box* box, void* stream) {
  if (int e = check_small_grid(g)) return e;
  if (num_p < 0 || num_p >= ((int64_t)1 << 31)) return fail(-1, "bad particle count %lld", (long long)num_p);
  if (num_p == 0) return 0;
  if (!pos || !spd || !box) return fail(-1, "null particle buffer or seed box");
  AdvectArgs a{MacView{g->mac, mac_is_half(g->form)}, g->cls, nullptr, nullptr, reinterpret_cast<float4*>(pos),
               spd, (uint32_t)num_p, g->nx, g->ny, g->nz, dt, uin, max_age, seed, *box};
  hipLaunchKernelGGL(k_advect, dim3((unsigned)((num_p + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0,
                     (hipStream_t)stream, a);
  return check_launch("wt_advect");
}

int wt_advect_ordered(const wt_grid* g, float* pos, float* spd, const uint32_t* order, int64_t num_p, float dt,
                      float uin, float max_age, uint32_t seed, const wt_seed_box* box, void* stream) {
  if (int e = check_small_grid(g)) return e;
  if (num_p < 0 || num_p >= ((int64_t)1 << 31)) return fail(-1, "bad particle count %lld", (long long)num_p);
  if (num_p == 0) return 0;
  if (!pos || !spd || !box || !order) return fail(-1, "null particle buffer, order or seed box");
  AdvectArgs a{MacView{g->mac, mac_is_half(g->form)}, g->cls, order, nullptr, reinterpret_cast<float4*>(pos),
               spd, (uint32_t)num_p, g->nx, g->ny, g->nz, dt, uin, max_age, seed, *box};
  hipLaunchKernelGGL(k_advect, dim3((unsigned)((num_p + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0,
                     (hipStream_t)stream, a);
  return check_launch("wt_advect_ordered");
}

int wt_advect_slots(const wt_grid* g, float* pos, float* spd, const uint32_t* ids, int64_t num_p, float dt,
                    float uin, float max_age, uint32_t seed, const wt_seed_box* box, void* stream) {
  if (int e = check_small_grid(g)) return e;
  if (num_p < 0 || num_p >= ((int64_t)1 << 31)) return fail(-1, "bad particle count %lld", (long long)num_p);
  if (num_p == 0) return 0;
  if (!pos || !spd || !box || !ids) return fail(-1, "null particle buffer, ids or seed box");
  AdvectArgs a{MacView{g->mac, mac_is_half(g->form)}, g->cls, nullptr, ids, reinterpret_cast<float4*>(pos),
               spd, (uint32_t)num_p, g->nx, g->ny, g->nz, dt, uin, max_age, seed, *box};
  hipLaunchKernelGGL(k_advect, dim3((unsigned)((num_p + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0,
                     (hipStream_t)stream, a);
  return check_launch("wt_advect_slots");
}

int wt_particle_gather(const float* pos, const float* spd, const uint32_t* ids, const uint32_t* order,
                       int64_t num_p, float* pos_out, float* spd_out, uint32_t* ids_out, void* stream) {
  if (num_p < 0 || num_p >= ((int64_t)1 << 31)) return fail(-1, "bad particle count %lld", (long long)num_p);
  if (num_p == 0) return 0;
  if (!pos || !spd || !order || !pos_out || !spd_out || !ids_out) return fail(-1, "null particle buffer");
  if (pos_out == pos || spd_out == spd || (ids && ids_out == ids))
    return fail(-1, "wt_particle_gather cannot work in place");
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((num_p + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0,
                     (hipStream_t)stream, reinterpret_cast<const float4*>(pos), spd, ids, order, (uint32_t)num_p,
                     reinterpret_cast<float4*>(pos_out), spd_out, ids_out);
  return check_launch("wt_particle_gather");
}

// Workspace: keys, sorted keys, ids (3 x num_p uint32, each 256-byte aligned), then radix-sort scratch.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int64_t wt_particle_order_workspace_bytes(int64_t num_p) {
  if (num_p <= 0 || num_p >= ((int64_t)1 << 31)) {
    fail(-1, "bad particle count %lld", (long long)num_p);
    return -1;
  }
  size_t tmp = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                         (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)num_p) != hipSuccess) {
    fail(-2, "radix sort size query failed");
    return -2;
  }
  return (int64_t)(3 * align256(4 * (size_t)num_p) + align256(tmp));
}

int wt_particle_order(const wt_grid* g, const float* pos, int64_t num_p, uint32_t* order, void* workspace,
                      int64_t workspace_bytes, void* stream) {
  if (int e = check_small_grid(g)) return e;
  if (!pos || !order || !workspace) return fail(-1, "null particle, order or workspace buffer");
  const int64_t need = wt_particle_order_workspace_bytes(num_p);
  if (need < 0) return (int)need;
  if (workspace_bytes < need) return fail(-1, "workspace of %lld bytes, need %lld", (long long)workspace_bytes,
                                          (long long)need);
  hipStream_t s = (hipStream_t)stream;
  const size_t a = align256(4 * (size_t)num_p);
  char* w = (char*)workspace;
  uint32_t* keys = (uint32_t*)w;
  uint32_t* keys_sorted = (uint32_t*)(w + a);
  uint32_t* ids = (uint32_t*)(w + 2 * a);
  void* tmp = w + 3 * a;
  size_t tmp_bytes = (size_t)(workspace_bytes - (int64_t)(3 * a));
  hipLaunchKernelGGL(k_particle_keys, dim3((unsigned)((num_p + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0, s,
                     reinterpret_cast<const float4*>(pos), keys, ids, (uint32_t)num_p, g->nx, g->ny, g->nz);
  if (int e = check_launch("wt_particle_order")) return e;
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)g->nx * g->ny * g->nz) bits++;
  if (hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_sorted, ids, order, (int)num_p, 0, bits, s) !=
      hipSuccess)
    return fail(-2, "radix sort failed: %s", hipGetErrorString(hipGetLastError()));
  return check_launch("wt_particle_order");
}

int wt_pack_particles(const wt_grid* g, const float* pos, const float* spd, uint16_t* q16, uint8_t* q8,
                      int64_t num_p, void* stream) {
  if (!g) return fail(-1, "null grid");
  if (num_p < 0 || num_p >= ((int64_t)1 << 31)) return fail(-1, "bad particle count %lld", (long long)num_p);
  if (num_p == 0) return 0;
  if (!pos || !spd || !q16 || !q8) return fail(-1, "null particle buffer");
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((num_p + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0,
                     (hipStream_t)stream, reinterpret_cast<const float4*>(pos), spd, nullptr,
                     reinterpret_cast<ushort4*>(q16), q8, (uint32_t)num_p, g->nx, g->ny, g->nz);
  return check_launch("wt_pack_particles");
}

int wt_pack_particles_slots(const wt_grid* g, const float* pos, const float* spd, const uint32_t* ids,
                            uint16_t* q16, uint8_t* q8, int64_t num_p, void* stream) {
  if (!g) return fail(-1, "null grid");
  if (num_p < 0 || num_p >= ((int64_t)1 << 31)) return fail(-1, "bad particle count %lld", (long long)num_p);
  if (num_p == 0) return 0;
  if (!pos || !spd || !ids || !q16 || !q8) return fail(-1, "null particle buffer or ids");
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((num_p + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0,
                     (hipStream_t)stream, reinterpret_cast<const float4*>(pos), spd, ids,
                     reinterpret_cast<ushort4*>(q16), q8, (uint32_t)num_p, g->nx, g->ny, g->nz);
  return check_launch("wt_pack_particles_slots");
}

int wt_slice(const wt_grid* g, uint8_t* out, int axis, int index, int field, float uin, float lref, float vmin,
             float vmax, void* stream) {
  if (int e = check_small_grid(g)) return e;
  if (!out) return fail(-1, "null output");
  if (axis < 0 || axis > 2) return fail(-1, "slice axis must be 0, 1 or 2 (got %d)", axis);
  const int dims[3] = {g->nx, g->ny, g->nz};
  if (index < 0 || index >= dims[axis])
    return fail(-1, "slice index %d outside [0, %d) on axis %d", index, dims[axis], axis);
  // slice_dims, lbm.py:455-456
  const int w = axis == 2 ? g->nx : (axis == 1 ? g->nx : g->nz);
  const int h = axis == 2 ? g->ny : (axis == 1 ? g->nz : g->ny);
  SliceArgs a{MacView{g->mac, mac_is_half(g->form)}, g->cls, out, g->nx, g->ny, g->nz, axis, index, field,
              w, (uint32_t)w * (uint32_t)h, uin, lref, vmin, vmax};
  hipLaunchKernelGGL(k_slice, dim3((a.npix + ABLOCK - 1) / ABLOCK), dim3(ABLOCK), 0, (hipStream_t)stream, a);
  return check_launch("wt_slice");
}

int wt_cp(const wt_grid* g, const float* verts, const float* norms, int8_t* out, int64_t nv, float uin,
          float cp_range, void* stream) {
  if (int e = check_small_grid(g)) return e;
  if (nv < 0 || nv >= ((int64_t)1 << 31)) return fail(-1, "bad vertex count %lld", (long long)nv);
  if (nv == 0) return 0;
  if (!verts || !norms || !out) return fail(-1, "null vertex, normal or output buffer");
  CpArgs a{MacView{g->mac, mac_is_half(g->form)}, g->cls, verts, norms, out, (uint32_t)nv,
           g->nx, g->ny, g->nz, uin, cp_range};
  hipLaunchKernelGGL(k_cp, dim3((unsigned)((nv + ABLOCK - 1) / ABLOCK)), dim3(ABLOCK), 0, (hipStream_t)stream, a);
  return check_launch("wt_cp");
}

}  // extern "C"
