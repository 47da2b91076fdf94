// Fused D3Q19 step for gfx950 and the kernels that share its cell logic.
//
// One launch = one lattice step of one tunnel (or of every tunnel of a batch):
// for each cell, in registers, it performs
//   face equilibrium boundary conditions (lbm.py:84-109),
//   pull streaming with bounce-back and the moving belt (lbm.py:110-124),
//   moments with the velocity clamp (lbm.py:125-136),
//   regularized collision with Smagorinsky LES and the outlet sponge
//   (lbm.py:137-165), and the macroscopic store (lbm.py:166),
// optionally followed by the momentum-exchange force (lbm.py:169-197) and the
// mass sum of the cells it just updated, reduced in FP64 with wavefront
// shuffles into fixed slots so the result does not depend on scheduling.
//
// Layout: populations are row-blocked SoA (wt_device.h: each x-row holds its 19
// planes in turn). A workgroup is 256 threads = 4 wave64; lanes run along x, so
// each population load or store of a wave is one contiguous 256 B (FP32) or
// 128 B (FP16) run, and the 19 stores of a row land in one region (38 KiB for
// an FP32 row of 512 cells) instead of 19 planes n cells apart.
// Link classes come from the class byte built once per geometry: only cells
// flagged as touching a solid gather their 18 neighbour classes.
//
// Face cells read the previous step's macroscopic values of their nearest
// interior cell from a small ping-pong snapshot of the six layers next to the
// faces (deterministic Jacobi rule; the reference reads mac[] in the same
// launch that writes it, lbm.py:87 vs lbm.py:166).
//
// k_step_adv launches a step together with the particle advection of the
// previous step (advect_kernel, lbm.py:233-263, shared code in wt_particles.h):
// workgroups spread through its grid advect from the macroscopic buffer that
// step wrote, the rest run the step's tiles into the other buffer.
//
// The whole library is compiled with -ffp-contract=off: each operation is
// rounded once, in the order the reference's generated source uses, so the
// FP32 path can be compared bit for bit with the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "wt_device.h"
#include "wt_host.h"
#include "wt_particles.h"

namespace wt {

constexpr int BLOCK = 256;       // threads per workgroup (4 wave64)
// The thread index as the step path reads it. The persistent batched step's unit
// (wt_step_persist.hip) reads it through an empty asm at every use, so that the
// compiler cannot hoist what depends on it out of the item loop.
#ifdef WT_STEP_PERSIST_UNIT
__device__ __forceinline__ uint32_t opaque_tid() {
  uint32_t t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}
#define WT_TID opaque_tid()
#else
#define WT_TID ((uint32_t)threadIdx.x)
#endif
constexpr int SLOT = 256;        // cells per reduction slot
constexpr int CHUNK = 1024;      // slots per first-level reduction block

template <typename T> __device__ __forceinline__ T clampv(T v, T a, T b) {
  return (v < a) ? a : ((v > b) ? b : v);
}
// wp.max(a, b) on floats, with the same choice for equal values as the oracle
template <typename T> __device__ __forceinline__ T maxv(T a, T b) { return (a < b) ? b : a; }

// Cache policy of the population streams, measured on MI355X at 33.5M cells with
// the row-blocked layout: non-temporal stores are 7-8 % faster than cached ones
// (FP32 and FP16; with plain SoA planes they had measured the other way round).
// Non-temporal loads of all planes are 9 % slower (the x +- 1 pulls of
// neighbouring waves share lines); of the c_x = 0 planes alone, whose aligned
// lines one wave reads whole, 0.9 % (FP32) and 0.5 % (FP16) faster per session
// step. Storing the macroscopic field non-temporally made session steps (step,
// particle advection and slices, which read the field) 0.8 % faster for FP16 and
// left FP32 the same. Single tunnels whose state partly fits the Infinity Cache
// lose with both, so of the single-tunnel launches only those with more than
// 2 GiB of populations use them (run_tile's POL 1, solo_policy); batched steps
// use them for FP32 (0.610 against 0.614-0.634 ms for the 48-tunnel batch) and
// not for the 2-byte forms (1 % slower), batch_policy. Both apply on the
// raw-buffer paths; the 64-bit global-address path loads every plane cached. A
// tunnel whose populations fit in 224 MiB stores them cached too (POL 2).
template <typename E> __device__ __forceinline__ E ld_stream(const E* p) { return *p; }
template <bool NT, typename E> __device__ __forceinline__ void st_stream(E* p, E v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
// The macroscopic field of cell idx in its storage form's layout (MacView):
// float4 (ux, uy, uz, rho), or for FP16/BF16 binary16 (ux, uy, uz, rho - 1).
template <int FORM, bool NT>
__device__ __forceinline__ void st_mac(void* mac, uint32_t idx, float a, float b, float c, float d) {
  if constexpr (Form<FORM>::dev) {
    const uint32_t lo = (uint32_t)f2h(a) | ((uint32_t)f2h(b) << 16);
    const uint32_t hi = (uint32_t)f2h(c) | ((uint32_t)f2h(d - 1.0f) << 16);
    u32x2* p = reinterpret_cast<u32x2*>(mac) + idx;
    if constexpr (NT) __builtin_nontemporal_store(u32x2{lo, hi}, p);
    else *p = u32x2{lo, hi};
  } else {
    f32x4* p = reinterpret_cast<f32x4*>(mac) + idx;
    if constexpr (NT) __builtin_nontemporal_store(f32x4{a, b, c, d}, p);
    else *p = f32x4{a, b, c, d};
  }
}

// Signed sum c_i . u in index order, as cu_expr() emits it (lbm.py:48-53).
template <typename R>
__device__ __forceinline__ R cdot(int i, R ux, R uy, R uz) {
  R s = (R)0;
  if (cx(i) > 0) s = s + ux; else if (cx(i) < 0) s = s - ux;
  if (cy(i) > 0) s = s + uy; else if (cy(i) < 0) s = s - uy;
  if (cz(i) > 0) s = s + uz; else if (cz(i) < 0) s = s - uz;
  return s;
}

// ---- shell (face-adjacent snapshot) ---------------------------------------------
template <typename R> struct Shell {
  R* base;
  int32_t nx, ny, nz;
  __device__ R* xlo() const { return base; }
  __device__ R* xhi() const { return base + 4 * (size_t)ny * nz; }
  __device__ R* ylo() const { return base + 8 * (size_t)ny * nz; }
  __device__ R* yhi() const { return base + 8 * (size_t)ny * nz + 4 * (size_t)nx * nz; }
  __device__ R* zlo() const { return base + 8 * ((size_t)ny * nz + (size_t)nx * nz); }
  __device__ R* zhi() const { return base + 8 * ((size_t)ny * nz + (size_t)nx * nz) + 4 * (size_t)nx * ny; }
};

template <typename R>
__device__ __forceinline__ void put4(R* p, R a, R b, R c, R d) {
  p[0] = a; p[1] = b; p[2] = c; p[3] = d;
}
template <>
__device__ __forceinline__ void put4<float>(float* p, float a, float b, float c, float d) {
  *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
}

// Interior and solid cells on one of the six layers publish their macroscopic
// values for the next step's face cells.
template <typename R>
__device__ __forceinline__ void shell_write(const Shell<R>& s, int x, int y, int z, R ux, R uy, R uz,
                                            R rho) {
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  if (x == 1) put4(s.xlo() + 4 * (size_t)(y + ny * z), ux, uy, uz, rho);
  if (x == nx - 2) put4(s.xhi() + 4 * (size_t)(y + ny * z), ux, uy, uz, rho);
  if (y == 1) put4(s.ylo() + 4 * (size_t)(x + nx * z), ux, uy, uz, rho);
  if (y == ny - 2) put4(s.yhi() + 4 * (size_t)(x + nx * z), ux, uy, uz, rho);
  if (z == 1) put4(s.zlo() + 4 * (size_t)(x + nx * y), ux, uy, uz, rho);
  if (z == nz - 2) put4(s.zhi() + 4 * (size_t)(x + nx * y), ux, uy, uz, rho);
}

// The previous-step values of a face cell's nearest interior cell
// clamp(x,1,nx-2), clamp(y,1,ny-2), clamp(z,1,nz-2) (lbm.py:87). Never called
// for x == 0, whose inlet values do not depend on it (lbm.py:96-100).
template <typename R>
__device__ __forceinline__ const R* shell_read(const Shell<R>& s, int x, int y, int z) {
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int tx = clampv(x, 1, nx - 2), ty = clampv(y, 1, ny - 2), tz = clampv(z, 1, nz - 2);
  if (x == nx - 1) return s.xhi() + 4 * (size_t)(ty + ny * tz);
  if (y == 0) return s.ylo() + 4 * (size_t)(tx + nx * tz);
  if (y == ny - 1) return s.yhi() + 4 * (size_t)(tx + nx * tz);
  if (z == 0) return s.zlo() + 4 * (size_t)(tx + nx * ty);
  return s.zhi() + 4 * (size_t)(tx + nx * ty);
}

// ---- link classification ------------------------------------------------------
// Bounce-back and belt bitmasks of an interior fluid cell from the classes of
// its 18 upstream neighbours idx - c_i (lbm.py:114-124).
struct Links {
  uint32_t bb, belt;
};
__device__ __forceinline__ int32_t offs(int i, int32_t nx, int32_t nxy) {
  return cx(i) + nx * cy(i) + nxy * cz(i);
}
__device__ __forceinline__ Links gather_links(const uint8_t* cls, uint32_t idx, int32_t nx,
                                              int32_t nxy) {
  Links l{0u, 0u};
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const uint32_t t = cls[idx - offs(i, nx, nxy)] & CLS_CODE;
    l.bb |= (t != 0u ? 1u : 0u) << i;
    if (cx(i) != 0) l.belt |= (t == 2u ? 1u : 0u) << i;
  }
  return l;
}
// Links i whose downstream neighbour idx + c_i is a static solid (lbm.py:188).
__device__ __forceinline__ uint32_t gather_force_links(const uint8_t* cls, uint32_t idx, int32_t nx,
                                                       int32_t nxy) {
  uint32_t m = 0;
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const uint32_t t = cls[idx + offs(i, nx, nxy)] & CLS_CODE;
    m |= (t == 1u ? 1u : 0u) << i;
  }
  return m;
}

// Link classes staged in LDS. A wave whose cells touch a wall (any lane flagged
// CLS_NEAR) copies the class bytes of the 9 rows its cells' neighbours lie in --
// rows (dy, dz) in {-1, 0, 1}^2 around its 64 consecutive cells, one cell beyond
// each end, from the dword boundary below -- into its own slice of LDS with
// 2-3 coalesced dword loads per lane, then every lane reads its 18 upstream
// (lbm.py:114-124) and 18 downstream (lbm.py:185-188) neighbour classes from
// LDS instead of gathering 36 single bytes from memory. The loads go through a
// raw buffer of the class array's n bytes: a window reaching before the first
// or past the last cell reads zeros there (offsets below 0 wrap above n), and
// only lanes of interior cells, whose neighbours all lie inside, use them.
// A wave reads back only what it wrote itself, so no workgroup barrier is needed.
constexpr uint32_t LINK_DW = 18;  // dwords per row window: 66 bytes from a dword boundary
__device__ __forceinline__ uint32_t row_start(uint32_t idx0, int w, int32_t nx, int32_t nxy) {
  const int dy = w / 3 - 1, dz = w % 3 - 1;
  return idx0 + (uint32_t)(dy * nx + dz * nxy) - 1u;
}
__device__ __forceinline__ void stage_link_rows(uint32_t (*rows)[LINK_DW], const uint8_t* cls, uint32_t n,
                                                uint32_t idx0, int32_t nx, int32_t nxy) {
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)cls, (short)0, (int)n, 0x00020000);
  const uint32_t lane = WT_TID & 63;
  uint32_t* flat = &rows[0][0];            // dword q is row q / LINK_DW, dword q % LINK_DW
  auto offset = [&](uint32_t q) {
    const uint32_t w = q / LINK_DW, k = q - w * LINK_DW;
    return (row_start(idx0, (int)w, nx, nxy) & ~3u) + 4u * k;
  };
  // the active lanes are a prefix of the wave (a tunnel's last wave may be partial)
  const uint32_t m = (uint32_t)__popcll(__ballot(1));
  if (m == 64) {
    // a full wave: the 162 dwords as 3 loads per lane in flight together
    const uint32_t q2 = lane + 128;
    const uint32_t v0 = __builtin_amdgcn_raw_buffer_load_b32(r, offset(lane), 0, 0);
    const uint32_t v1 = __builtin_amdgcn_raw_buffer_load_b32(r, offset(lane + 64), 0, 0);
    const uint32_t v2 = q2 < 9 * LINK_DW ? __builtin_amdgcn_raw_buffer_load_b32(r, offset(q2), 0, 0) : 0u;
    flat[lane] = v0;
    flat[lane + 64] = v1;
    if (q2 < 9 * LINK_DW) flat[q2] = v2;
  } else {
    for (uint32_t q = lane; q < 9 * LINK_DW; q += m) flat[q] = __builtin_amdgcn_raw_buffer_load_b32(r, offset(q), 0, 0);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// class code of the neighbour (dx, dy, dz) of the lane's cell, from the staged rows
__device__ __forceinline__ uint32_t staged_code(const uint32_t (*rows)[LINK_DW], uint32_t idx0, int32_t nx,
                                                int32_t nxy, int dx, int dy, int dz) {
  const int w = (dy + 1) * 3 + (dz + 1);
  const uint32_t p = (row_start(idx0, w, nx, nxy) & 3u) + (WT_TID & 63) + 1u + (uint32_t)dx;
  return reinterpret_cast<const uint8_t*>(rows[w])[p] & CLS_CODE;
}
__device__ __forceinline__ Links staged_links(const uint32_t (*rows)[LINK_DW], uint32_t idx0, int32_t nx,
                                              int32_t nxy) {
  Links l{0u, 0u};
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const uint32_t t = staged_code(rows, idx0, nx, nxy, -cx(i), -cy(i), -cz(i));
    l.bb |= (t != 0u ? 1u : 0u) << i;
    if (cx(i) != 0) l.belt |= (t == 2u ? 1u : 0u) << i;
  }
  return l;
}
__device__ __forceinline__ uint32_t staged_force_links(const uint32_t (*rows)[LINK_DW], uint32_t idx0, int32_t nx,
                                                       int32_t nxy) {
  uint32_t m = 0;
#pragma unroll
  for (int i = 1; i < Q; i++) m |= (staged_code(rows, idx0, nx, nxy, cx(i), cy(i), cz(i)) == 1u ? 1u : 0u) << i;
  return m;
}

struct Acc {
  double fx, fy, fz, m;
};

// ---- collision (lbm.py:125-165) --------------------------------------------------------
// Velocity clamp |u| <= 0.35 (lbm.py:130-135); returns ux^2 + uy^2 + uz^2 of the
// clamped velocity (usq = 1.5 times it, lbm.py:136). The square root is taken only
// where ux^2 + uy^2 + uz^2 > 0.12: correctly rounded sqrt is monotonic and
// sqrt(0.12) rounds to 0.3464 < 0.35, so below that no lane can clamp and the
// result equals the unconditional test of the reference bit for bit (NaN fails
// both comparisons alike).
template <typename R>
__device__ __forceinline__ R clamp_velocity(R& ux, R& uy, R& uz) {
  R u2 = ux * ux + uy * uy + uz * uz;
  if (u2 > (R)0.12) {
    const R um = sqrt(u2);
    if (um > (R)0.35) {
      const R sc = (R)0.35 / um;
      ux = ux * sc; uy = uy * sc; uz = uz * sc;
      u2 = ux * ux + uy * uy + uz * uz;
    }
  }
  return u2;
}

// Signed sum over the directions with a nonzero coefficient k(i), in index order,
// starting from the first such term (whose coefficient is +1 for every moment
// used below): the deviation form's sums, which carry no leading 0.0.
template <typename R, typename K>
__device__ __forceinline__ R signed_sum(const R (&v)[Q], K k) {
  R acc = (R)0;
  bool any = false;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const int c = k(i);
    if (c == 0) continue;
    if (!any) { acc = c > 0 ? v[i] : -v[i]; any = true; }
    else if (c > 0) acc = acc + v[i];
    else acc = acc - v[i];
  }
  return acc;
}

// Deviation-form collision (FP16 and BF16 storage): the same regularized
// Smagorinsky collision as collide() below, carried out on g_i = f_i - w_i in
// FP32, so no weight is added after the load or subtracted before the store.
//   dr = sum g_i, rho = 1 + dr, (jx, jy, jz) = sum c_i g_i, u = j / rho (clamped);
//   excess equilibrium eps_i = e_i - w_i = w_i dr + (w_i rho)(cu + (h - usq)),
//   cu = 3 c_i.u, h = (0.5 cu) cu;  d_i = g_i - eps_i;  Pi = sum c c d;
//   g_out_i = eps_i + 4.5 w_i keep (c_i c_i : Pi - tr Pi / 3).
// Opposite directions share h and give cu_{i+1} = -cu_i exactly, so the pair
// form below equals a per-direction evaluation. The oracle restates this
// sequence operation by operation for these forms.
__device__ __forceinline__ void collide_dev(float (&g)[Q], int x, int nx, float tau, float smag, float (&out)[Q],
                                            float& ux, float& uy, float& uz, float& rho) {
  float dr = g[0];
#pragma unroll
  for (int i = 1; i < Q; i++) dr = dr + g[i];
  rho = 1.0f + dr;
  const float ir = 1.0f / rho;
  const float jx = signed_sum(g, [](int i) { return cx(i); });
  const float jy = signed_sum(g, [](int i) { return cy(i); });
  const float jz = signed_sum(g, [](int i) { return cz(i); });
  ux = jx * ir; uy = jy * ir; uz = jz * ir;
  const float usq = 1.5f * clamp_velocity<float>(ux, uy, uz);
  const float w0 = (float)wd(0), w1 = (float)wd(1), w2 = (float)wd(7);
  const float a0 = w0 * dr, a1 = w1 * dr, a2 = w2 * dr;
  const float b0 = w0 * rho, b1 = w1 * rho, b2 = w2 * rho;
  float eps[Q];
  eps[0] = a0 - b0 * usq;
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    // c_i for odd i has a first nonzero component of +1
    float sdot;
    if (cx(i) != 0) {
      sdot = ux;
      if (cy(i) > 0) sdot = sdot + uy; else if (cy(i) < 0) sdot = sdot - uy;
      if (cz(i) > 0) sdot = sdot + uz; else if (cz(i) < 0) sdot = sdot - uz;
    } else if (cy(i) != 0) {
      sdot = uy;
      if (cz(i) > 0) sdot = sdot + uz; else if (cz(i) < 0) sdot = sdot - uz;
    } else {
      sdot = uz;
    }
    const float cu = 3.0f * sdot;
    const float h = 0.5f * cu * cu;
    const float hu = h - usq;
    const float a = i <= 6 ? a1 : a2, b = i <= 6 ? b1 : b2;
    eps[i] = a + b * (cu + hu);
    eps[i + 1] = a + b * (hu - cu);
  }
  float d[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) d[i] = g[i] - eps[i];
  const float pxx = signed_sum(d, [](int i) { return cx(i) * cx(i); });
  const float pyy = signed_sum(d, [](int i) { return cy(i) * cy(i); });
  const float pzz = signed_sum(d, [](int i) { return cz(i) * cz(i); });
  const float pxy = signed_sum(d, [](int i) { return cx(i) * cy(i); });
  const float pxz = signed_sum(d, [](int i) { return cx(i) * cz(i); });
  const float pyz = signed_sum(d, [](int i) { return cy(i) * cz(i); });
  const float qn = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
  float tau0 = tau;
  const float sp = (float)x - (float)nx + 24.0f;
  if (sp > 0.0f) tau0 = tau0 + sp * 0.015f;
  const float tau_e = 0.5f * (tau0 + sqrt(tau0 * tau0 + 25.4558441f * smag * smag * qn * ir));
  const float keep = 1.0f - 1.0f / tau_e;
  const float tr3 = (pxx + pyy + pzz) * (1.0f / 3.0f);
  const float k0 = (float)(4.5 * wd(0)) * keep, k1 = (float)(4.5 * wd(1)) * keep, k2 = (float)(4.5 * wd(7)) * keep;
  out[0] = eps[0] - k0 * tr3;
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    float acc2 = 0.0f;
    bool any = false;
    if (cx(i) != 0) { acc2 = pxx; any = true; }
    if (cy(i) != 0) { acc2 = any ? acc2 + pyy : pyy; any = true; }
    if (cz(i) != 0) { acc2 = any ? acc2 + pzz : pzz; any = true; }
    if (cx(i) * cy(i) != 0) { const float t = (float)(2 * cx(i) * cy(i)) * pxy; acc2 = any ? acc2 + t : t; any = true; }
    if (cx(i) * cz(i) != 0) { const float t = (float)(2 * cx(i) * cz(i)) * pxz; acc2 = any ? acc2 + t : t; any = true; }
    if (cy(i) * cz(i) != 0) { const float t = (float)(2 * cy(i) * cz(i)) * pyz; acc2 = any ? acc2 + t : t; any = true; }
    const float gk = (i <= 6 ? k1 : k2) * (acc2 - tr3);
    out[i] = eps[i] + gk;
    out[i + 1] = eps[i + 1] + gk;
  }
}

// Collision of a storage form: deviation form for FP16/BF16, the reference's
// order of operations (collide) for FP32 and FP64.
template <int FORM>
__device__ __forceinline__ void collide_form(typename Form<FORM>::real (&f)[Q], int x, int nx,
                                             typename Form<FORM>::real tau, typename Form<FORM>::real smag,
                                             typename Form<FORM>::real (&out)[Q], typename Form<FORM>::real& ux,
                                             typename Form<FORM>::real& uy, typename Form<FORM>::real& uz,
                                             typename Form<FORM>::real& rho);
// Moments, velocity clamp, regularized collision with Smagorinsky LES and the
// outlet sponge, for populations f after streaming; writes post-collision
// populations to out. Opposite directions (i, i+1 for odd i) share their
// equilibrium terms: c_{i+1} = -c_i makes cu_{i+1} = -cu_i and
// (0.5 cu)cu identical bit for bit, and the second-order projection identical,
// so each shared value is exactly what a per-direction evaluation computes.
template <typename R>
__device__ __forceinline__ void collide(R (&f)[Q], int x, int nx, R tau, R smag, R (&out)[Q], R& ux, R& uy,
                                        R& uz, R& rho) {
  const R zero = (R)0, one = (R)1;
  rho = zero;
#pragma unroll
  for (int i = 0; i < Q; i++) rho = rho + f[i];
  const R ir = one / rho;
  R sx = zero, sy = zero, sz = zero;
#pragma unroll
  for (int i = 0; i < Q; i++) { if (cx(i) > 0) sx = sx + f[i]; else if (cx(i) < 0) sx = sx - f[i]; }
#pragma unroll
  for (int i = 0; i < Q; i++) { if (cy(i) > 0) sy = sy + f[i]; else if (cy(i) < 0) sy = sy - f[i]; }
#pragma unroll
  for (int i = 0; i < Q; i++) { if (cz(i) > 0) sz = sz + f[i]; else if (cz(i) < 0) sz = sz - f[i]; }
  ux = sx * ir; uy = sy * ir; uz = sz * ir;
  const R usq = (R)1.5 * clamp_velocity<R>(ux, uy, uz);
  const R wr0 = (R)wd(0) * rho, wr1 = (R)wd(1) * rho, wr2 = (R)wd(7) * rho;
  // equilibria: e_i = (w rho) * (((1 + cu) + (0.5 cu) cu) - usq)
  R e[Q];
  e[0] = wr0 * ((one + zero + (R)0.5 * zero * zero) - usq);
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    const R cu = (R)3.0 * cdot<R>(i, ux, uy, uz);
    const R h = (R)0.5 * cu * cu;
    const R wr = i <= 6 ? wr1 : wr2;
    e[i] = wr * ((one + cu + h) - usq);
    e[i + 1] = wr * ((one + (-cu) + h) - usq);
  }
  // non-equilibrium second moments, each summed in index order (lbm.py:141-146)
  R pxx = zero, pyy = zero, pzz = zero, pxy = zero, pxz = zero, pyz = zero;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const R di = f[i] - e[i];
    if (cx(i) != 0) pxx = pxx + di;
    if (cy(i) != 0) pyy = pyy + di;
    if (cz(i) != 0) pzz = pzz + di;
    if (cx(i) * cy(i) > 0) pxy = pxy + di; else if (cx(i) * cy(i) < 0) pxy = pxy - di;
    if (cx(i) * cz(i) > 0) pxz = pxz + di; else if (cx(i) * cz(i) < 0) pxz = pxz - di;
    if (cy(i) * cz(i) > 0) pyz = pyz + di; else if (cy(i) * cz(i) < 0) pyz = pyz - di;
  }
  // Smagorinsky relaxation time and outlet sponge (lbm.py:147-154)
  const R qn = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + (R)2.0 * (pxy * pxy + pxz * pxz + pyz * pyz));
  R tau0 = tau;
  const R sp = (R)x - (R)nx + (R)24.0;
  if (sp > zero) tau0 = tau0 + sp * (R)0.015;
  const R tau_e = (R)0.5 * (tau0 + sqrt(tau0 * tau0 + (R)25.4558441 * smag * smag * qn * ir));
  const R keep = one - one / tau_e;
  const R tr3 = (pxx + pyy + pzz) * ((R)1.0 / (R)3.0);
  const R k0 = (R)(4.5 * wd(0)) * keep, k1 = (R)(4.5 * wd(1)) * keep, k2 = (R)(4.5 * wd(7)) * keep;
  // regularized collision (lbm.py:155-165): f = e + (4.5 w keep) (c c : Pi - tr Pi / 3)
  out[0] = e[0] + k0 * (zero - tr3);
#pragma unroll
  for (int i = 1; i < Q; i += 2) {
    R acc2 = zero;
    bool any = false;
    if (cx(i) != 0) { acc2 = pxx; any = true; }
    if (cy(i) != 0) { acc2 = any ? acc2 + pyy : pyy; any = true; }
    if (cz(i) != 0) { acc2 = any ? acc2 + pzz : pzz; any = true; }
    if (cx(i) * cy(i) != 0) { const R t = (R)(2 * cx(i) * cy(i)) * pxy; acc2 = any ? acc2 + t : t; any = true; }
    if (cx(i) * cz(i) != 0) { const R t = (R)(2 * cx(i) * cz(i)) * pxz; acc2 = any ? acc2 + t : t; any = true; }
    if (cy(i) * cz(i) != 0) { const R t = (R)(2 * cy(i) * cz(i)) * pyz; acc2 = any ? acc2 + t : t; any = true; }
    const R g = (i <= 6 ? k1 : k2) * (acc2 - tr3);
    out[i] = e[i] + g;
    out[i + 1] = e[i + 1] + g;
  }
}

template <int FORM>
__device__ __forceinline__ void collide_form(typename Form<FORM>::real (&f)[Q], int x, int nx,
                                             typename Form<FORM>::real tau, typename Form<FORM>::real smag,
                                             typename Form<FORM>::real (&out)[Q], typename Form<FORM>::real& ux,
                                             typename Form<FORM>::real& uy, typename Form<FORM>::real& uz,
                                             typename Form<FORM>::real& rho) {
  if constexpr (Form<FORM>::dev) collide_dev(f, x, nx, tau, smag, out, ux, uy, uz, rho);
  else collide<typename Form<FORM>::real>(f, x, nx, tau, smag, out, ux, uy, uz, rho);
}

// Momentum exchange of one interior cell on the populations it stores
// (lbm.py:185-193): links i whose downstream neighbour is a static solid.
template <int FORM>
__device__ __forceinline__ void add_force(uint32_t fl, const typename Form<FORM>::elem (&st)[Q], Acc& acc) {
  typedef Form<FORM> S;
  typedef typename S::real R;
#pragma unroll
  for (int i = 1; i < Q; i++) {
    if ((fl >> i) & 1u) {
      const R v = (R)2.0 * S::excess(st[i], i);
      const double dv = (double)v;
      if (cx(i) > 0) acc.fx += dv; else if (cx(i) < 0) acc.fx -= dv;
      if (cy(i) > 0) acc.fy += dv; else if (cy(i) < 0) acc.fy -= dv;
      if (cz(i) > 0) acc.fz += dv; else if (cz(i) < 0) acc.fz -= dv;
    }
  }
}

// ---- face equilibrium boundaries (lbm.py:84-109) -------------------------------------
// Boundary density and velocity of a face cell: inlet values at x == 0, else the
// previous step's values of the nearest interior cell with the wall and outlet
// overrides.
// m: the snapshot values of the nearest interior cell (unused at x == 0).
template <typename R>
__device__ __forceinline__ void face_state_of(const StepDesc& d, const R* m, int x, int y, int z,
                                              R& br, R& bux, R& buy, R& buz) {
  const R zero = (R)0, one = (R)1;
  if (x == 0) {
    br = one; bux = (R)d.uin; buy = zero; buz = zero;
    return;
  }
  br = clampv(m[3], (R)0.9, (R)1.1);
  bux = m[0]; buy = m[1]; buz = m[2];
  if (y == 0 || y == d.ny - 1) buy = zero;
  if (z == 0 || z == d.nz - 1) buz = zero;
  if (x == d.nx - 1) { br = one; bux = maxv(bux, zero); }
}
template <typename R>
__device__ __forceinline__ void face_state(const StepDesc& d, const Shell<R>& shin, int x, int y, int z,
                                           R& br, R& bux, R& buy, R& buz) {
  if (x == 0) face_state_of<R>(d, nullptr, x, y, z, br, bux, buy, buz);
  else face_state_of<R>(d, shell_read<R>(shin, x, y, z), x, y, z, br, bux, buy, buz);
}

// Equilibrium populations of a face cell (lbm.py:104-107).
template <typename R>
__device__ __forceinline__ R face_feq(int i, R br, R bux, R buy, R buz, R bsq) {
  const R bc = (R)3.0 * cdot<R>(i, bux, buy, buz);
  return (R)wd(i) * br * ((R)1 + bc + (R)0.5 * bc * bc - bsq);
}

// ---- population accesses ------------------------------------------------------------
// For 2- and 4-byte storage the populations are addressed as one raw buffer
// (hardware range check on) with 32-bit offsets; its base sits `bias` bytes
// below the allocation so that every pull offset is non-negative. Buffers
// beyond 32-bit byte offsets (and FP64) use 64-bit global addresses instead.
constexpr int BUF_GFX9_WORD3 = 0x00020000;  // raw buffer, gfx9-family resource word 3
// Directions in ascending memory order of their upstream rows (z-1 rows first,
// then y-1, the own row, y+1, z+1; x+1 before x-1 within a row): issuing the 19
// loads in this order measured 0.5-0.7 % faster than direction order.
constexpr int LOAD_ORDER[Q] = {15, 5, 11, 14, 18, 3, 7, 10, 0, 1, 2, 4, 8, 9, 17, 6, 12, 13, 16};
// cache-policy word of the raw-buffer loads and stores (gfx950: bit 0 sc0, bit 1 nt, bit 4 sc1)
constexpr int CPOL_LD = 0;      // population loads, cached
constexpr int CPOL_ST = 2;      // non-temporal population stores
constexpr int CPOL_LD_NT = 2;   // loads of the c_x = 0 planes

template <typename E> struct RawIO;
template <> struct RawIO<float> {
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, bool nt) {
    return __uint_as_float(nt ? __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, CPOL_LD_NT)
                              : __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, CPOL_LD));
  }
  __device__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v, bool nt) {
    if (nt) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, vo, so, CPOL_ST);
    else __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, vo, so, 0);
  }
};
template <> struct RawIO<uint16_t> {
  __device__ static uint16_t ld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, bool nt) {
    return nt ? __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, CPOL_LD_NT)
              : __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, CPOL_LD);
  }
  __device__ static void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, uint16_t v, bool nt) {
    if (nt) __builtin_amdgcn_raw_buffer_store_b16(v, r, vo, so, CPOL_ST);
    else __builtin_amdgcn_raw_buffer_store_b16(v, r, vo, so, 0);
  }
};

// A hook the step paths call once per cell after the collision and before the
// stores, on every lane that steps a cell (at that point a wave has no store
// outstanding from this cell). The persistent batched step publishes its next
// claim there; every other kernel passes NoHook.
struct NoHook {
  __device__ __forceinline__ void operator()() const {}
};

// ---- one cell, general path (any mix of solid, face and near-wall lanes) ------------
template <int FORM, int POL, class H = NoHook>
__device__ __forceinline__ void step_cell_general(const StepDesc& d, const typename Form<FORM>::elem* fin,
                                                  typename Form<FORM>::elem* fout,
                                                  const Shell<typename Form<FORM>::real>& shin,
                                                  const Shell<typename Form<FORM>::real>& shout,
                                                  uint32_t idx, uint32_t pc, uint8_t c, int x, int y, int z,
                                                  bool face, int flags, Acc& acc, const H& hook = H()) {
  typedef Form<FORM> S;
  typedef typename S::real R;
  typedef typename S::elem E;
  const int32_t nx = d.nx, ny = d.ny;
  const int32_t nxy = nx * ny;
  const R zero = (R)0, one = (R)1;
  // class bytes of the wave's neighbour rows, when any lane touches a wall
  __shared__ uint32_t link_rows[BLOCK / 64][9][LINK_DW];
  uint32_t (*rows)[LINK_DW] = link_rows[WT_TID >> 6];
  const uint32_t idx0 = idx - (WT_TID & 63);
  if (__any((c & CLS_NEAR) != 0)) stage_link_rows(rows, d.cls, d.n, idx0, nx, nxy);

  const bool solid = (c & CLS_CODE) != 0;  // solid cell, lbm.py:81-83
  constexpr uint32_t ES = (uint32_t)sizeof(E);
  const uint32_t n = d.n;
  const uint32_t bo = pc * ES;
  // stores of the 19 populations of this cell (own row block, plane offsets)
  auto store_all = [&](const E (&v)[Q]) {
    auto global_stores = [&]() {
#pragma unroll
      for (int i = 0; i < Q; i++) st_stream<POL != 2>(fout + (pc + (uint32_t)ppos(i) * (uint32_t)nx), v[i]);
    };
    if constexpr (sizeof(E) <= 4) {
      if (!d.global_addr) {
        const __amdgpu_buffer_rsrc_t rout =
            __builtin_amdgcn_make_buffer_rsrc((void*)fout, (short)0, (int)((uint32_t)Q * n * ES), BUF_GFX9_WORD3);
#pragma unroll
        for (int i = 0; i < Q; i++) RawIO<E>::st(rout, bo, (uint32_t)ppos(i) * (uint32_t)nx * ES, v[i], POL != 2);
      } else {
        global_stores();
      }
    } else {
      global_stores();
    }
  };
  // Face lanes fetch their snapshot values first and take their boundary state
  // after the interior lanes have issued their loads and collided, as in the fast
  // path below (the other order measured 4.3 % (FP32) and 8.4 % (FP16) slower on
  // the 48-tunnel batch, whose small tunnels put face lanes in many general waves,
  // and 0.4-0.6 % slower at 33.5M cells).
  typedef R R4 __attribute__((ext_vector_type(4)));
  R4 sm = __builtin_nondeterministic_value(R4{});
  if (face && !solid && x != 0) sm = *reinterpret_cast<const R4*>(shell_read<R>(shin, x, y, z));
  R out[Q], ux, uy, uz, rho;
  if (!solid && !face) {
    // pull streaming with bounce-back and the moving belt, lbm.py:110-124:
    // population i comes from element pc + pshift(i) (the upstream neighbour), or
    // from the cell's own opposite plane pc + ppos(opp(i)) * nx when that
    // neighbour is a wall; the loads go out in ascending upstream-row order
    Links lk{0u, 0u};
    if (c & CLS_NEAR) lk = staged_links(rows, idx0, nx, nxy);
    R f[Q];
    auto global_loads = [&]() {
#pragma unroll
      for (int k = 0; k < Q; k++) {
        const int i = LOAD_ORDER[k];
        const bool b = (lk.bb >> i) & 1u;
        const uint32_t a = pc + (b ? (uint32_t)ppos(opp(i)) * (uint32_t)nx : pshift(i, nx, nxy));
        f[i] = S::pop(ld_stream(fin + a));
      }
    };
    if constexpr (sizeof(E) <= 4) {
      if (!d.global_addr) {
        // per-lane offset bo + bias + (element shift) * ES: non-negative, and below
        // Q * n * ES + bias for every element of the array (no 32-bit wrap)
        const uint32_t bias = (uint32_t)Q * ((uint32_t)nxy + (uint32_t)nx) * ES;
        const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
            (void*)((const char*)fin - bias), (short)0, (int)((uint32_t)Q * n * ES + bias), BUF_GFX9_WORD3);
#pragma unroll
        for (int k = 0; k < Q; k++) {
          const int i = LOAD_ORDER[k];
          // branch-free select between the two wave-uniform offsets
          const uint32_t pull = pshift(i, nx, nxy) * ES + bias;
          const uint32_t wall = (uint32_t)ppos(opp(i)) * (uint32_t)nx * ES + bias;
          const uint32_t m = 0u - ((lk.bb >> i) & 1u);
          f[i] = S::pop(RawIO<E>::ld(rin, bo + (pull + (m & (wall - pull))), 0u, POL == 1 && cx(i) == 0));
        }
      } else {
        global_loads();
      }
    } else {
      global_loads();
    }
    if (lk.belt) {
      const R uin = (R)d.uin;
#pragma unroll
      for (int i = 1; i < Q; i++)
        if (cx(i) != 0 && ((lk.belt >> i) & 1u)) f[i] = f[i] + (R)(6.0 * wd(i) * cx(i)) * uin;
    }
    collide_form<FORM>(f, x, nx, (R)d.tau, (R)d.smag, out, ux, uy, uz, rho);
  }
  asm volatile("" : "+v"(sm));
  hook();
  if (!solid && face) {  // equilibrium boundaries, lbm.py:84-109
    const R m[4] = {sm.x, sm.y, sm.z, sm.w};
    face_state_of<R>(d, m, x, y, z, rho, ux, uy, uz);
    const R bsq = (R)1.5 * (ux * ux + uy * uy + uz * uz);
    // the deviation forms store g = f - w, as store() does
#pragma unroll
    for (int i = 0; i < Q; i++)
      out[i] = S::dev ? face_feq<R>(i, rho, ux, uy, uz, bsq) - (R)wd(i) : face_feq<R>(i, rho, ux, uy, uz, bsq);
  }
  if (solid) {
    if (!(flags & WT_FLAG_NO_MAC)) st_mac<FORM, POL == 1>(d.mac, idx, 0.f, 0.f, 0.f, 1.f);
    shell_write<R>(shout, x, y, z, zero, zero, zero, one);
    return;
  }
  E st[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) st[i] = S::put(out[i]);
  store_all(st);
  if (!(flags & WT_FLAG_NO_MAC)) st_mac<FORM, POL == 1>(d.mac, idx, (float)ux, (float)uy, (float)uz, (float)rho);
  if (!face) shell_write<R>(shout, x, y, z, ux, uy, uz, rho);
  if (flags & WT_FLAG_MASS) acc.m += (double)rho;
  if ((flags & WT_FLAG_FORCES) && !face && (c & CLS_FORCE))
    add_force<FORM>(staged_force_links(rows, idx0, nx, nxy), st, acc);
}

// ---- one cell, fast path: the whole wave is interior fluid away from walls -----------
// Every lane pulls from its 18 neighbours. With the row-blocked layout the load
// of population i is the lane's own element offset (one vector register shared
// by all 38 loads and stores) plus the wave-uniform pshift(i) in a scalar
// register (raw-buffer offsets, as above), with no per-lane address arithmetic.
template <int FORM, int POL, class H = NoHook>
__device__ __forceinline__ void step_cell_fast(const StepDesc& d, const typename Form<FORM>::elem* fin,
                                               typename Form<FORM>::elem* fout,
                                               const Shell<typename Form<FORM>::real>& shin,
                                               const Shell<typename Form<FORM>::real>& shout, uint32_t idx,
                                               uint32_t pc, int x, int y, int z, bool xface, int flags,
                                               Acc& acc, const H& hook = H()) {
  typedef Form<FORM> S;
  typedef typename S::real R;
  typedef typename S::elem E;
  const uint32_t nx = (uint32_t)d.nx, n = d.n, nxy = nx * (uint32_t)d.ny;
  constexpr uint32_t ES = (uint32_t)sizeof(E);
  const uint32_t bo = pc * ES;
  const uint32_t bias = (uint32_t)Q * (nxy + nx) * ES;
  R out[Q], ux, uy, uz, rho;
  // Inlet / outlet lanes (lbm.py:84-109) take their boundary state after the
  // interior lanes' collision, with the outlet lanes' snapshot values fetched
  // before the wave's population loads go out: the other order held those waves'
  // loads back behind the snapshot load (measured: the 48-tunnel batch 0.8 % and
  // 128x64x64 tunnels 2.4-3.6 % faster, 33.5M cells the same, bit-identical).
  // One vector, left unset on the other lanes: a split or zero-filled copy would
  // make the compiler wait for the load at the join.
  typedef R R4 __attribute__((ext_vector_type(4)));
  R4 sm = __builtin_nondeterministic_value(R4{});
  if (xface && x != 0) sm = *reinterpret_cast<const R4*>(shell_read<R>(shin, x, y, z));
  if (!xface) {
    R f[Q];
    auto global_loads = [&]() {
#pragma unroll
      for (int k = 0; k < Q; k++) {
        const int i = LOAD_ORDER[k];
        f[i] = S::pop(ld_stream(fin + (pc + pshift(i, nx, nxy))));
      }
    };
    if constexpr (sizeof(E) <= 4) {
      if (!d.global_addr) {
        const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
            (void*)((const char*)fin - bias), (short)0, (int)((uint32_t)Q * n * ES + bias), BUF_GFX9_WORD3);
#pragma unroll
        for (int k = 0; k < Q; k++) {
          const int i = LOAD_ORDER[k];
          f[i] = S::pop(RawIO<E>::ld(rin, bo, pshift(i, nx, nxy) * ES + bias, POL == 1 && cx(i) == 0));
        }
      } else {
        global_loads();
      }
    } else {
      global_loads();
    }
    collide_form<FORM>(f, x, nx, (R)d.tau, (R)d.smag, out, ux, uy, uz, rho);
  }
  // the boundary arithmetic may not move above this point (the compiler would
  // otherwise hoist it next to the snapshot loads, and wait for them there)
  asm volatile("" : "+v"(sm));
  hook();
  if (xface) {
    // inlet / outlet lanes of this wave (lbm.py:84-109)
    const R m[4] = {sm.x, sm.y, sm.z, sm.w};
    face_state_of<R>(d, m, x, y, z, rho, ux, uy, uz);
    const R bsq = (R)1.5 * (ux * ux + uy * uy + uz * uz);
    // in the deviation forms the common store below takes g = f - w: the same
    // subtraction store() makes on the face path
#pragma unroll
    for (int i = 0; i < Q; i++)
      out[i] = S::dev ? face_feq<R>(i, rho, ux, uy, uz, bsq) - (R)wd(i) : face_feq<R>(i, rho, ux, uy, uz, bsq);
  }
  auto global_stores = [&]() {
#pragma unroll
    for (int i = 0; i < Q; i++) st_stream<POL != 2>(fout + (pc + (uint32_t)ppos(i) * nx), S::put(out[i]));
  };
  if constexpr (sizeof(E) <= 4) {
    if (!d.global_addr) {
      const __amdgpu_buffer_rsrc_t rout =
          __builtin_amdgcn_make_buffer_rsrc((void*)fout, (short)0, (int)((uint32_t)Q * n * ES), BUF_GFX9_WORD3);
#pragma unroll
      for (int i = 0; i < Q; i++) RawIO<E>::st(rout, bo, (uint32_t)ppos(i) * nx * ES, S::put(out[i]), POL != 2);
    } else {
      global_stores();
    }
  } else {
    global_stores();
  }
  if (!(flags & WT_FLAG_NO_MAC)) st_mac<FORM, POL == 1>(d.mac, idx, (float)ux, (float)uy, (float)uz, (float)rho);
  if (!xface) shell_write<R>(shout, x, y, z, ux, uy, uz, rho);
  if (flags & WT_FLAG_MASS) acc.m += (double)rho;
}

template <int FORM, int POL, class H = NoHook>
__device__ __forceinline__ void step_cell(const StepDesc& d, const typename Form<FORM>::elem* fin,
                                          typename Form<FORM>::elem* fout,
                                          const Shell<typename Form<FORM>::real>& shin,
                                          const Shell<typename Form<FORM>::real>& shout,
                                          uint32_t idx, int flags, Acc& acc, const H& hook = H()) {
  const int32_t nx = d.nx, ny = d.ny, nz = d.nz;
  const uint32_t yz = divu(idx, d.mx, d.sx);
  const int x = (int)(idx - yz * (uint32_t)nx);
  const uint32_t zz = divu(yz, d.my, d.sy);
  const int y = (int)(yz - zz * (uint32_t)ny);
  const int z = (int)zz;
  const uint32_t pc = prow(idx, yz, (uint32_t)nx);
  const uint8_t c = d.cls[idx];
  const bool xface = (x == 0) | (x == nx - 1);
  const bool face = xface | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1);
  // fluid lanes away from walls, and fluid inlet/outlet lanes, take the fast path
  const bool simple = (c == 0) & (!face | xface) & (y != 0) & (y != ny - 1) & (z != 0) & (z != nz - 1);
  if (__all(simple))
    step_cell_fast<FORM, POL, H>(d, fin, fout, shin, shout, idx, pc, x, y, z, xface, flags, acc, hook);
  else
    step_cell_general<FORM, POL, H>(d, fin, fout, shin, shout, idx, pc, c, x, y, z, face, flags, acc, hook);
}

// ---- deterministic FP64 reduction ------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Each 256-cell slot is reduced the same way whatever the launch shape: the
// 64 lanes of each of its 4 waves by the shuffle tree above, then the 4 wave
// sums in wave order.
__device__ __forceinline__ void slot_stage(double (*lds)[4][4], int k, Acc a) {
  const double fx = wave_sum(a.fx), fy = wave_sum(a.fy), fz = wave_sum(a.fz), m = wave_sum(a.m);
  const int lane = WT_TID & 63, wave = WT_TID >> 6;
  if (lane == 0) {
    lds[k][wave][0] = fx; lds[k][wave][1] = fy; lds[k][wave][2] = fz; lds[k][wave][3] = m;
  }
}

template <int CPT>
__device__ __forceinline__ void slot_flush(double (*lds)[4][4], double* partials, uint32_t slot0,
                                           uint32_t nslots) {
  __syncthreads();
  const int t = WT_TID;
  if (t < 4 * CPT) {
    const int k = t >> 2, comp = t & 3;
    if (slot0 + k < nslots) {
      const double s = ((lds[k][0][comp] + lds[k][1][comp]) + lds[k][2][comp]) + lds[k][3][comp];
      partials[4 * (size_t)(slot0 + k) + comp] = s;
    }
  }
}

// ---- the tile loop ------------------------------------------------------------------
// POL, the cache policy of a launch (chosen by cache_policy on the host):
//   0  FP16/BF16/FP64 batches and mid-size tunnels: populations stored
//      non-temporally, the rest cached;
//   1  FP32 batches and one tunnel with more than 2 GiB of populations: also the
//      macroscopic field stored and the c_x = 0 planes loaded non-temporally
//      (33.5M cells: session steps 0.9-1.05 % faster for FP32, 0.7-1.2 % for FP16;
//      the 48-tunnel FP32 batch 1.3-1.5 % faster, FP16 1 % slower; smaller single
//      tunnels lose with it);
//   2  one tunnel whose two population buffers fit in 224 MiB: everything cached,
//      so the Infinity Cache carries a step's output into the next step.
// NBUF slot staging areas in LDS, area lbuf for this tile: the persistent
// batched step alternates two, so that the next tile's staging needs no barrier
// after this tile's flush.
template <int FORM, int CPT, int POL, int NBUF = 1, class H = NoHook>
__device__ __forceinline__ void run_tile(const StepDesc& d, int parity, uint32_t tile, int flags, int lbuf = 0,
                                         const H& hook = H()) {
  typedef Form<FORM> S;
  typedef typename S::real R;
  typedef typename S::elem E;
  __shared__ double slots[NBUF][CPT][4][4];
  double (*lds)[4][4] = slots[NBUF > 1 ? lbuf : 0];
  const int side = parity & 1;
  const E* fin = (const E*)(side ? d.fout : d.fin);
  E* fout = (E*)(side ? d.fin : d.fout);
  const Shell<R> shin{(R*)(side ? d.shell_out : d.shell_in), d.nx, d.ny, d.nz};
  const Shell<R> shout{(R*)(side ? d.shell_in : d.shell_out), d.nx, d.ny, d.nz};
  const uint32_t base = tile * (uint32_t)(BLOCK * CPT);
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    const uint32_t idx = base + (uint32_t)(k * BLOCK) + WT_TID;
    Acc a{0.0, 0.0, 0.0, 0.0};
    if (idx < d.n) step_cell<FORM, POL, H>(d, fin, fout, shin, shout, idx, flags, a, hook);
    if (flags & (WT_FLAG_FORCES | WT_FLAG_MASS)) slot_stage(lds, k, a);
  }
  if (flags & (WT_FLAG_FORCES | WT_FLAG_MASS))
    slot_flush<CPT>(lds, d.partials, tile * CPT, (d.n + SLOT - 1) / SLOT);
}

// Tile visiting order. Bit 0: workgroups are dispatched to the 8 XCDs
// round-robin; map them so each XCD walks one contiguous range of tiles
// (neighbouring tiles share cache lines through the x +- 1 pulls). Bit 1: walk
// the tiles backwards, so that a step started right after a forward step
// first reads what that step wrote last, while it is still in the caches.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblocks, int remap) {
  uint32_t t = b;
  if (remap & 1) {
    const uint32_t q = nblocks >> 3, r = nblocks & 7u;
    const uint32_t xcd = b & 7u, k = b >> 3;
    t = xcd * q + (xcd < r ? xcd : r) + k;
  }
  return (remap & 2) ? nblocks - 1u - t : t;
}

template <int FORM, int CPT, int POL>
__global__ __launch_bounds__(BLOCK) void k_step(StepDesc d, int parity, int remap, int flags) {
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x, remap);
  run_tile<FORM, CPT, POL>(d, parity, tile, flags);
}

// One lattice step plus the particle advection of the previous step from its
// macroscopic field a.mac (a buffer the step does not write), in one launch.
// The nadv = gridDim.x - step_blocks advection workgroups are spread through
// the grid: with s = gridDim.x / nadv, workgroup q s advects particle block q
// (q < nadv) and every other workgroup runs the next tile in the tile order of
// k_step, so the latency-bound gathers run beside the HBM-bound tiles
// throughout. Measured with the FP16 session step at 33.5M cells and 2^20
// particles (two boxes, 3 and 4 alternating runs): the fused kernel's median
// 0.536-0.542 ms spread, 0.539-0.545 ms with the advection taking the first
// workgroups, 0.547-0.549 ms after the tiles. (An earlier spread over the
// kernels of the time had measured slower than advection first.)
// Every workgroup maps to exactly one tile or particle block. The tile order's
// XCD bit (xcd_tile, bit 0) sees the tile's rank k among the step workgroups,
// not the workgroup index, so in this grid a tile's XCD follows the spread
// rather than that order: only the speed of the shapes that set the bit
// (tunnels below 200k cells) can change, never a result.
template <int FORM, int CPT, int POL>
__global__ __launch_bounds__(BLOCK) void k_step_adv(StepDesc d, AdvectArgs a, int parity, int remap,
                                                                 int flags, uint32_t step_blocks) {
  const uint32_t nadv = gridDim.x - step_blocks;
  const uint32_t b = blockIdx.x;
  bool adv = false;
  uint32_t k = b;  // tile index, or particle block when adv
  if (nadv > 0) {
    const uint32_t s = gridDim.x / nadv;
    if (b < nadv * s) {
      const uint32_t q = b / s, r = b - q * s;
      adv = r == 0;
      k = adv ? q : q * (s - 1) + r - 1;
    } else {
      k = nadv * (s - 1) + (b - nadv * s);
    }
  }
  if (!adv) {
    run_tile<FORM, CPT, POL>(d, parity, xcd_tile(k, step_blocks, remap), flags);
  } else {
    const uint32_t j = k * BLOCK + threadIdx.x;
    if (j < a.num_p) advect_particle(a, j);
  }
}


// Batched step: one workgroup per work item, tile t of tunnel k for every tunnel
// of the batch, tunnel k's tiles holding 256 * ds[k].cpt cells (the launch shape
// of its size class, one unrolled code path per cells-per-thread value up to
// MAXCPT, the batch's largest: the variant without the 2- and 4-cell paths
// needs 64-65 registers instead of 80; FP16 on 48 tunnels 10.7 % faster). With
// `backwards` the items are dispatched from the last to the first. Each cell is
// computed exactly as by k_step. The persistent form, k_step_batch_persist
// below, runs 1.7-3.9 % (FP32) and 8-10 % (FP16) slower on 48 tunnels and is
// not the default. Its first version, a grid of one occupancy striding over the
// items in this unit, measured 16 % slower: the loop raised the kernel to 96
// registers, 5 waves per SIMD instead of 7; on the one-cell variant it spilled
// 21-74 scalar registers, 106-115 VGPRs for FP32, 25 VGPRs to scratch for FP16.
// Those register measures are why it now has a unit of its own.
// Waves per SIMD the compiler fits the batched step to: 8 for the FP16/BF16
// one-cell variant (64 registers, no spills; measured 1.6 % faster than its
// natural 7), no constraint otherwise.
template <int FORM, int MAXCPT> constexpr int batch_min_waves() { return Form<FORM>::dev && MAXCPT == 1 ? 8 : 1; }
// Cache policy of the batched step (run_tile's POL): 1 for FP32 (the 48-tunnel
// batch 1.3-1.5 % faster than with 0, the serving window the same), 0 otherwise
// (FP16: 1 % slower with 1; the all-cached 2: FP32 13.6 %, FP16 1.7 % slower).
template <int FORM> constexpr int batch_policy() { return FORM == WT_FP32 ? 1 : 0; }

// gfx950's constant address space (scalar loads of memory no launch writes); the
// host pass of this single-source file sees plain pointers
#ifdef __HIP_DEVICE_COMPILE__
#define WT_CONSTANT_AS __attribute__((address_space(4)))
#else
#define WT_CONSTANT_AS
#endif

// One work item of a batched step: finds the item's tunnel and runs its tile.
template <int FORM, int MAXCPT, int NBUF = 1, class H = NoHook>
__device__ __forceinline__ void batch_item(const StepDesc* __restrict__ ds, const uint8_t* __restrict__ gates,
                                          int nd, uint32_t item, int parity, int flags, int lbuf = 0,
                                          const H& hook = H()) {
  // the descriptor table and gate bytes are written before the launch and only read
  // by it: reading them through the constant address space lets the compiler keep
  // the item's descriptor in scalar registers inside the persistent kernel's loop too
  typedef const WT_CONSTANT_AS StepDesc* CDesc;
  typedef const WT_CONSTANT_AS uint32_t* CWord;
  const CDesc cds = (CDesc)ds;
  // the item's tunnel: for FP16/BF16 the number of tunnels whose first tile is at
  // or before it, less one (ds[0].tile_begin is 0), counted 64 tunnels per
  // wave-wide vote: one round of loads instead of a chain of dependent ones
  // (48 tunnels: 2.4 % faster); for FP32/FP64 a binary search, whose later start
  // of each workgroup's streams measured faster there (the vote: 3.5 % slower)
  int lo = 0;
  if constexpr (Form<FORM>::dev) {
    lo = -1;
    for (int c0 = 0; c0 < nd; c0 += 64) {
      const int k = c0 + (int)(WT_TID & 63);
      lo += __popcll(__ballot(k < nd && cds[k].tile_begin <= item));
    }
  } else {
    int hi = nd - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cds[mid].tile_begin <= item) lo = mid; else hi = mid - 1;
    }
  }
  const StepDesc d = cds[lo];
  const int p = parity ^ (d.flags & 1);
  const uint32_t tile = item - d.tile_begin;
  // with WT_FLAG_NO_MAC a tunnel whose gate is set still stores its macroscopic field
  // (read as the gate's aligned word, a scalar load: the gate bytes are padded to a
  // multiple of 8 and start 8-byte aligned in the workspace)
  const uint32_t gw = ((CWord)gates)[lo >> 2];
  const int fl = ((flags & WT_FLAG_NO_MAC) && ((gw >> (8 * (lo & 3))) & 0xffu)) ? (flags & ~WT_FLAG_NO_MAC) : flags;
  if constexpr (MAXCPT >= 4) {
    if (d.cpt == 4) { run_tile<FORM, 4, batch_policy<FORM>(), NBUF, H>(d, p, tile, fl, lbuf, hook); return; }
  }
  if constexpr (MAXCPT >= 2) {
    if (d.cpt == 2) { run_tile<FORM, 2, batch_policy<FORM>(), NBUF, H>(d, p, tile, fl, lbuf, hook); return; }
  }
  run_tile<FORM, 1, batch_policy<FORM>(), NBUF, H>(d, p, tile, fl, lbuf, hook);
}

template <int FORM, int MAXCPT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(batch_min_waves<FORM, MAXCPT>())))
void k_step_batch(const StepDesc* __restrict__ ds, const uint8_t* __restrict__ gates, int nd, uint32_t nitems,
                  int parity, int backwards, int flags) {
  const uint32_t item = backwards ? nitems - 1u - blockIdx.x : blockIdx.x;
  batch_item<FORM, MAXCPT>(ds, gates, nd, item, parity, flags);
}

// Persistent batched step (WT_BATCH_PERSISTENT): a grid of the device's full
// occupancy whose workgroups claim the items dynamically, in order, from nctr
// interleaved counters (nctr a power of 2, at most 8 and at most the
// workgroups; workgroup b takes items k, k + nctr, k + 2 nctr, ... of counter
// k = b % nctr; workgroups are dispatched to the 8 XCDs round-robin, so with 8
// counters each XCD's workgroups share one). The items in flight then form one
// compact front through the batch, as with one workgroup per item (a static
// stride over the items, where workgroups drift apart, measured 19 % slower).
// A workgroup's first item is its own; it claims the next while it runs the
// current one. Every workgroup's last claim(s) fail (exceed its counter's
// items); the one that draws the last value of all resets the counter to 0 for
// the next launch, and the host encodes the counters as 0 in the first place, so
// the kernel can be captured and replayed. The arguments are re-read from the
// argument segment on every item, so that nothing but the claims is live across
// the item loop (with the unit's build options: the per-item code's registers).
constexpr int BATCH_COUNTERS = 8;
constexpr int BATCH_COUNTER_STRIDE = 32;  // words: one 128-byte line per counter
struct BatchArgs {
  const StepDesc* ds;
  const uint8_t* gates;
  uint32_t* counters;
  int32_t nd;
  uint32_t nitems;
  int32_t parity;
  int32_t backwards;
  int32_t flags;
  uint32_t nblocks;   // workgroups of the launch
  uint32_t log_nctr;  // counters in use: 2^log_nctr <= min(BATCH_COUNTERS, nblocks)
};

// The claims of an FP32 workgroup come from a fifth wave that runs no tile (CW):
// the compiler makes any use of an atomic's result wait for every store the wave
// has outstanding, so a tile wave taking the claims stalls on its own population
// stores. The other forms, which run more waves per SIMD, let thread 0 of the
// first tile wave claim.
// Measured on the 48-tunnel batch: FP32 with the claim wave 2 % slower than one
// workgroup per item, with thread 0 claiming 6 % slower; FP16 with thread 0 claiming
// 8-10 % slower, with the claim wave 22 % slower (its fifth wave takes a quarter of
// the tile waves' occupancy: 8 waves per SIMD at 64 registers).
constexpr int CLAIM_WAVE_THREADS = 64;
template <bool CW> constexpr int persist_block() { return CW ? BLOCK + CLAIM_WAVE_THREADS : BLOCK; }
template <int FORM> constexpr bool persist_claim_wave() { return FORM == WT_FP32; }

// Thread 0's claims, taken in the tile code's pre-store hook (the item's loads
// are in and its stores not yet out, so the claim's atomic neither holds the
// loads back nor waits for the stores): the hook publishes the claim that came
// back since the previous item's hook and sends the next one, once per item.
// It fires on every item: thread 0 always steps the tile's first cell (a tile
// starts below n), and both step paths call the hook outside any per-lane branch.
struct ClaimHook {
  uint32_t* slot;
  uint32_t* pend;
  bool* fired;
  uint32_t* ctr;
  uint32_t base;
  bool active;
  __device__ __forceinline__ void operator()() const {
    if (active && !*fired) {
      *slot = *pend;
      *pend = base + atomicAdd(ctr, 1u);
      *fired = true;
    }
  }
};

template <int FORM, int MAXCPT, bool CW>
__global__ __launch_bounds__(persist_block<CW>()) __attribute__((amdgpu_waves_per_eu(batch_min_waves<FORM, MAXCPT>())))
void k_step_batch_persist(BatchArgs args) {
  __shared__ uint32_t claim[2];
  // tile waves 0-3, the claim wave 4 (a wave-uniform value: each wave takes one side)
  const bool claimer = CW && __builtin_amdgcn_readfirstlane(WT_TID >> 6) == BLOCK / 64;
  const uint32_t claim_tid = CW ? BLOCK : 0;
  // a workgroup's first item is its own (j = b / nctr of counter b % nctr, one per
  // workgroup: there are at least as many items as workgroups); the claims then
  // hand out the rest of each counter's items in order (claim value v: j = wk + v).
  // Values are read into scalar registers: the item, its tunnel and its
  // descriptor are workgroup-uniform, and the tile code addresses through them.
  // Without the claim wave thread 0 keeps one claim in flight beyond the one it
  // publishes (sent at the start, then one per item from the hook), so each
  // workgroup's last two claims fail; with it, the claim wave sends one per item
  // and each workgroup's last claim fails.
  uint32_t j, pend = 0;
  {
    const WT_CONSTANT_AS BatchArgs* a0 = (const WT_CONSTANT_AS BatchArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    const uint32_t ln = a0->log_nctr, nc = 1u << ln, k = blockIdx.x & (nc - 1u);
    j = blockIdx.x >> ln;
    if (!CW && WT_TID == 0)
      pend = ((a0->nblocks - k + nc - 1u) >> ln) + atomicAdd(a0->counters + BATCH_COUNTER_STRIDE * k, 1u);
  }
  for (int it = 0;; it ^= 1) {
    const WT_CONSTANT_AS BatchArgs* a = (const WT_CONSTANT_AS BatchArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(a));
    const uint32_t ln = a->log_nctr, nc = 1u << ln, k = blockIdx.x & (nc - 1u), nitems = a->nitems;
    uint32_t* ctr = a->counters + BATCH_COUNTER_STRIDE * k;
    const uint32_t nk = (nitems - k + nc - 1u) >> ln;    // items k, k + nc, ... below nitems
    const uint32_t wk = (a->nblocks - k + nc - 1u) >> ln;  // workgroups sharing the counter
    if (j >= nk) {
      // the last claim of all (value nk - 1, resp. nk + wk - 1) resets the counter
      // for the next launch
      if (CW ? (WT_TID == claim_tid && j == nk + wk - 1u) : (WT_TID == 0 && pend == nk + 2u * wk - 1u))
        atomicExch(ctr, 0u);
      break;
    }
    const int flags = a->flags;
    // the next claim reaches LDS before the item's barrier (the tile's slot flush
    // when it reduces forces or mass, otherwise an explicit one); claim slots
    // alternate, so a slot is rewritten only after everyone has read it
    if (claimer) {
      if (WT_TID == claim_tid) claim[it] = wk + atomicAdd(ctr, 1u);
      __syncthreads();
    } else {
      bool fired = false;
      const ClaimHook hook{&claim[it], &pend, &fired, ctr, wk, !CW && WT_TID == 0};
      const uint32_t item = k + (j << ln);
      batch_item<FORM, MAXCPT, 2, ClaimHook>(a->ds, a->gates, a->nd, a->backwards ? nitems - 1u - item : item,
                                             a->parity, flags, it, hook);
      if (!(flags & (WT_FLAG_FORCES | WT_FLAG_MASS))) __syncthreads();
    }
    j = __builtin_amdgcn_readfirstlane(claim[it]);
  }
}

#if !defined(WT_STEP_FP32_UNIT) && !defined(WT_STEP_PERSIST_UNIT)  // kernels launched only from the main unit
// ---- reductions of the slots ---------------------------------------------------------
struct RedDesc {
  double* partials;  // nslots slots, then nchunks first-level sums
  double* sums;      // 4 doubles (step parity 0)
  double* sums_odd;  // 4 doubles (step parity 1)
  uint32_t nslots;
};

__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  return ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

// First level: block (c, t) sums slots [c*CHUNK, (c+1)*CHUNK) of tunnel t.
__global__ __launch_bounds__(BLOCK) void k_reduce1(const RedDesc* __restrict__ rd) {
  __shared__ double lds[4];
  const RedDesc r = rd[blockIdx.y];
  const uint32_t nchunks = (r.nslots + CHUNK - 1) / CHUNK;
  if (blockIdx.x >= nchunks) return;
  const uint32_t s0 = blockIdx.x * CHUNK + threadIdx.x * (CHUNK / BLOCK);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < CHUNK / BLOCK; j++) {
    const uint32_t s = s0 + j;
    if (s < r.nslots) {
#pragma unroll
      for (int c = 0; c < 4; c++) v[c] += r.partials[4 * (size_t)s + c];
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const double t = block_sum(v[c], lds);
    if (threadIdx.x == 0) r.partials[4 * ((size_t)r.nslots + blockIdx.x) + c] = t;
  }
}

// Second level: block t sums its first-level chunks in order.
__global__ __launch_bounds__(BLOCK) void k_reduce2(const RedDesc* __restrict__ rd, int parity) {
  __shared__ double lds[4];
  const RedDesc r = rd[blockIdx.x];
  const uint32_t nchunks = (r.nslots + CHUNK - 1) / CHUNK;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t ch = threadIdx.x; ch < nchunks; ch += BLOCK) {
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] += r.partials[4 * ((size_t)r.nslots + ch) + c];
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const double t = block_sum(v[c], lds);
    if (threadIdx.x == 0) (parity ? r.sums_odd : r.sums)[c] = t;
  }
}

// Single-tunnel variants take the descriptor by value.
__global__ __launch_bounds__(BLOCK) void k_reduce1_one(RedDesc r) {
  __shared__ double lds[4];
  const uint32_t s0 = blockIdx.x * CHUNK + threadIdx.x * (CHUNK / BLOCK);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < CHUNK / BLOCK; j++) {
    const uint32_t s = s0 + j;
    if (s < r.nslots) {
#pragma unroll
      for (int c = 0; c < 4; c++) v[c] += r.partials[4 * (size_t)s + c];
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const double t = block_sum(v[c], lds);
    if (threadIdx.x == 0) r.partials[4 * ((size_t)r.nslots + blockIdx.x) + c] = t;
  }
}
__global__ __launch_bounds__(BLOCK) void k_reduce2_one(RedDesc r) {
  __shared__ double lds[4];
  const uint32_t nchunks = (r.nslots + CHUNK - 1) / CHUNK;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t ch = threadIdx.x; ch < nchunks; ch += BLOCK) {
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] += r.partials[4 * ((size_t)r.nslots + ch) + c];
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const double t = block_sum(v[c], lds);
    if (threadIdx.x == 0) r.sums[c] = t;
  }
}

// ---- standalone force on f[cur] (lbm.py:169-197) ---------------------------------------
template <int FORM>
__global__ __launch_bounds__(BLOCK) void k_force(StepDesc d) {
  typedef Form<FORM> S;
  typedef typename S::real R;
  typedef typename S::elem E;
  __shared__ double lds[1][4][4];
  const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
  Acc a{0.0, 0.0, 0.0, 0.0};
  if (idx < d.n) {
    const uint8_t c = d.cls[idx];
    if (c & CLS_FORCE) {
      const E* f = (const E*)d.fin;
      const int32_t nx = d.nx, nxy = d.nx * d.ny;
      const uint32_t pc = prow(idx, divu(idx, d.mx, d.sx), (uint32_t)nx);
      const uint32_t fl = gather_force_links(d.cls, idx, nx, nxy);
#pragma unroll
      for (int i = 1; i < Q; i++) {
        if ((fl >> i) & 1u) {
          const R v = (R)2.0 * S::excess(f[pc + (uint32_t)ppos(i) * (uint32_t)nx], i);
          const double dv = (double)v;
          if (cx(i) > 0) a.fx += dv; else if (cx(i) < 0) a.fx -= dv;
          if (cy(i) > 0) a.fy += dv; else if (cy(i) < 0) a.fy -= dv;
          if (cz(i) > 0) a.fz += dv; else if (cz(i) < 0) a.fz -= dv;
        }
      }
    }
  }
  slot_stage(lds, 0, a);
  slot_flush<1>(lds, d.partials, blockIdx.x, (d.n + SLOT - 1) / SLOT);
}

// ---- init (lbm.py:64-69) ----------------------------------------------------------------
template <int FORM>
__global__ __launch_bounds__(BLOCK) void k_init(typename Form<FORM>::elem* f0,
                                                typename Form<FORM>::elem* f1, void* mac,
                                                uint32_t n, uint32_t nx, float uin_f) {
  typedef Form<FORM> S;
  typedef typename S::real R;
  const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= n) return;
  const uint32_t pc = prow(idx, idx / nx, nx);
  const R uin = (R)uin_f;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const R v = (R)wd(i) * ((R)1.0 + (R)(3 * cx(i)) * uin + (R)(4.5 * cx(i) * cx(i)) * uin * uin -
                            (R)1.5 * uin * uin);
    const typename S::elem e = S::store(v, i);
    f0[pc + (uint32_t)ppos(i) * nx] = e;
    f1[pc + (uint32_t)ppos(i) * nx] = e;
  }
  st_mac<FORM, false>(mac, idx, uin_f, 0.f, 0.f, 1.f);
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_init_shell(R* s0, R* s1, uint32_t entries, float uin) {
  const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= entries) return;
  put4<R>(s0 + 4 * (size_t)e, (R)uin, (R)0, (R)0, (R)1);
  put4<R>(s1 + 4 * (size_t)e, (R)uin, (R)0, (R)0, (R)1);
}

// ---- class bytes from the solid mask (lbm.py:419-421 input) --------------------------------
__global__ __launch_bounds__(BLOCK) void k_build_cls(const uint8_t* __restrict__ solid, uint8_t* cls,
                                                     int32_t nx, int32_t ny, int32_t nz, uint32_t n,
                                                     uint64_t mx, uint32_t sx, uint64_t my,
                                                     uint32_t sy) {
  const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= n) return;
  const uint32_t yz = divu(idx, mx, sx);
  const int x = (int)(idx - yz * (uint32_t)nx);
  const uint32_t zz = divu(yz, my, sy);
  const int y = (int)(yz - zz * (uint32_t)ny);
  const int z = (int)zz;
  auto code = [](uint8_t v) -> uint32_t { return v == 0 ? 0u : (v == 1 ? 1u : (v == 2 ? 2u : 3u)); };
  const uint32_t own = code(solid[idx]);
  uint32_t out = own;
  const bool face = (x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1);
  if (own == 0 && !face) {
    const int32_t nxy = nx * ny;
    bool near = false, force = false;
#pragma unroll
    for (int i = 1; i < Q; i++) {
      const uint32_t t = code(solid[idx + offs(i, nx, nxy)]);
      near |= (t != 0);
      force |= (t == 1);
    }
    if (near) out |= CLS_NEAR;
    if (force) out |= CLS_FORCE;
  }
  cls[idx] = (uint8_t)out;
}

// ---- link classes as the step kernel uses them --------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_link_classes(const uint8_t* __restrict__ cls, uint64_t* out,
                                                        int32_t nx, int32_t ny, int32_t nz, uint32_t n,
                                                        uint64_t mx, uint32_t sx, uint64_t my,
                                                        uint32_t sy) {
  const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= n) return;
  const uint32_t yz = divu(idx, mx, sx);
  const int x = (int)(idx - yz * (uint32_t)nx);
  const uint32_t zz = divu(yz, my, sy);
  const int y = (int)(yz - zz * (uint32_t)ny);
  const int z = (int)zz;
  const uint8_t c = cls[idx];
  // the step kernel's own path: neighbour classes staged in LDS per wave
  __shared__ uint32_t link_rows[BLOCK / 64][9][LINK_DW];
  uint32_t (*rows)[LINK_DW] = link_rows[threadIdx.x >> 6];
  const uint32_t idx0 = idx - (threadIdx.x & 63);
  const int32_t nxy = nx * ny;
  if (__any((c & CLS_NEAR) != 0)) stage_link_rows(rows, cls, n, idx0, nx, nxy);
  uint64_t v;
  if ((c & CLS_CODE) != 0) {
    v = 1;
  } else if ((x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1)) {
    v = 2;
  } else {
    Links lk{0u, 0u};
    if (c & CLS_NEAR) lk = staged_links(rows, idx0, nx, nxy);
    const uint32_t fl = (c & CLS_FORCE) ? staged_force_links(rows, idx0, nx, nxy) : 0u;
    v = ((uint64_t)lk.bb << 8) | ((uint64_t)lk.belt << 27) | ((uint64_t)fl << 45);
  }
  out[idx] = v;
}

#endif  // WT_STEP_FP32_UNIT

// ================================ host side =========================================

// Launch-shape choice for one storage form: cells per thread and XCD tile order.
struct Shape {
  int cpt;
  int remap;
  int solo;   // the cache policy of the launch (run_tile's POL: 0, 1 or 2)
};

// Workgroups of the single-tunnel step kernels reserve LDS they do not use,
// which holds them to fewer waves than their registers allow (160 KB of LDS per
// CU). FP32: under the streaming policy 1 (tunnels above 2 GiB of populations)
// 4 waves per SIMD with 36 KiB (33.5M cells: 1.8-2.0 % faster than 5; 3 waves
// 0.7 % faster than 5), otherwise 5 with 28 KiB (4.2M cells: 4 waves 1 % slower,
// 0.5M cells 3.5 % slower; 6 waves 1.7 % slower at 33.5M cells). FP16/BF16:
// under policy 1 6 waves with 22 KiB for the plain step (33.5M cells: 1.0-1.1 %
// faster than their natural 7, 5 waves 1 % slower), otherwise their 7, and their
// 7 for the step fused with the advection (the reserve holds the advection's
// workgroups back too: FP16 sessions 0.6 % slower with it).
constexpr size_t FP32_STEP_PAD_LDS = 28 * 1024;
constexpr size_t FP32_STREAM_STEP_PAD_LDS = 36 * 1024;
constexpr size_t HALF_STREAM_STEP_PAD_LDS = 22 * 1024;
template <int FORM, int POL> constexpr size_t step_pad_lds() {
  if (FORM == WT_FP32) return POL == 1 ? FP32_STREAM_STEP_PAD_LDS : FP32_STEP_PAD_LDS;
  if (Form<FORM>::dev) return POL == 1 ? HALF_STREAM_STEP_PAD_LDS : 0;
  return 0;
}
template <int FORM, int POL> constexpr size_t step_adv_pad_lds() {
  return FORM == WT_FP32 ? step_pad_lds<FORM, POL>() : 0;
}

template <int FORM, int CPT>
static void launch_step_t(const StepDesc& d, int parity, int remap, int flags, int solo, hipStream_t s) {
  const uint32_t ntile = (d.n + BLOCK * CPT - 1) / (BLOCK * CPT);
  switch (solo) {
    case 1: hipLaunchKernelGGL((k_step<FORM, CPT, 1>), dim3(ntile), dim3(BLOCK), (step_pad_lds<FORM, 1>()), s, d, parity, remap, flags); break;
    case 2: hipLaunchKernelGGL((k_step<FORM, CPT, 2>), dim3(ntile), dim3(BLOCK), (step_pad_lds<FORM, 2>()), s, d, parity, remap, flags); break;
    default: hipLaunchKernelGGL((k_step<FORM, CPT, 0>), dim3(ntile), dim3(BLOCK), (step_pad_lds<FORM, 0>()), s, d, parity, remap, flags); break;
  }
}

template <int FORM>
void launch_step_f(const StepDesc& d, int parity, Shape sh, int flags, hipStream_t s) {
  switch (sh.cpt) {
    case 2: launch_step_t<FORM, 2>(d, parity, sh.remap, flags, sh.solo, s); break;
    case 4: launch_step_t<FORM, 4>(d, parity, sh.remap, flags, sh.solo, s); break;
    default: launch_step_t<FORM, 1>(d, parity, sh.remap, flags, sh.solo, s); break;
  }
}

template <int FORM, int CPT>
static void launch_step_adv_t(const StepDesc& d, const AdvectArgs& a, int remap, int flags, int solo, hipStream_t s) {
  const uint32_t ntile = (d.n + BLOCK * CPT - 1) / (BLOCK * CPT);
  const uint32_t nadv = (a.num_p + BLOCK - 1) / BLOCK;
  switch (solo) {
    case 1: hipLaunchKernelGGL((k_step_adv<FORM, CPT, 1>), dim3(ntile + nadv), dim3(BLOCK), (step_adv_pad_lds<FORM, 1>()), s, d, a, 0, remap, flags, ntile); break;
    case 2: hipLaunchKernelGGL((k_step_adv<FORM, CPT, 2>), dim3(ntile + nadv), dim3(BLOCK), (step_adv_pad_lds<FORM, 2>()), s, d, a, 0, remap, flags, ntile); break;
    default: hipLaunchKernelGGL((k_step_adv<FORM, CPT, 0>), dim3(ntile + nadv), dim3(BLOCK), (step_adv_pad_lds<FORM, 0>()), s, d, a, 0, remap, flags, ntile); break;
  }
}
template <int FORM>
void launch_step_adv_f(const StepDesc& d, const AdvectArgs& a, Shape sh, int flags, hipStream_t s) {
  switch (sh.cpt) {
    case 2: launch_step_adv_t<FORM, 2>(d, a, sh.remap, flags, sh.solo, s); break;
    case 4: launch_step_adv_t<FORM, 4>(d, a, sh.remap, flags, sh.solo, s); break;
    default: launch_step_adv_t<FORM, 1>(d, a, sh.remap, flags, sh.solo, s); break;
  }
}

// The batched FP32 step (streaming cache policy) runs best at 5 waves per SIMD
// (48 tunnels, same box: 6 waves 0.4-0.8 % slower, 7 waves 1.7 %, 4 waves 0.8 %);
// 28 KiB of LDS per workgroup holds it there whatever its register count.
// FP16/BF16 take their 8 waves (64 registers).
constexpr size_t FP32_BATCH_PAD_LDS = 28 * 1024;
template <int FORM> constexpr size_t batch_pad_lds() { return FORM == WT_FP32 ? FP32_BATCH_PAD_LDS : 0; }
// Workgroups of a persistent launch: as many as the device holds at once (the
// kernel's occupancy per CU times the CUs), found once per device and kernel.
template <typename K>
static uint32_t resident_blocks(K kernel, int block, size_t lds, std::atomic<int>* cache, uint32_t nitems) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nitems;
  int g = cache[dev].load(std::memory_order_relaxed);
  if (g <= 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu <= 0 ||
        cus <= 0)
      return nitems;
    g = per_cu * cus;
    cache[dev].store(g, std::memory_order_relaxed);
  }
  return (uint32_t)g < nitems ? (uint32_t)g : nitems;
}

// Persistent batched step: a grid of one full occupancy, each workgroup striding
// over the items (instantiated in wt_step_persist.hip).
template <int FORM, int MAXCPT, bool CW>
static void launch_batch_persist_t(const StepDesc* dd, const uint8_t* gates, uint32_t* counters, int nd,
                                   uint32_t nitems, int parity, int backwards, int flags, uint32_t cap,
                                   hipStream_t s) {
  static std::atomic<int> cache[64];
  const size_t pad = batch_pad_lds<FORM>();
  uint32_t g = resident_blocks(k_step_batch_persist<FORM, MAXCPT, CW>, persist_block<CW>(), pad, cache, nitems);
  if (cap > 0 && cap < g) g = cap;
  uint32_t log_nctr = 0;
  while ((2u << log_nctr) <= g && (2u << log_nctr) <= (uint32_t)BATCH_COUNTERS) log_nctr++;
  const BatchArgs a{dd, gates, counters, nd, nitems, parity, backwards, flags, g, log_nctr};
  hipLaunchKernelGGL((k_step_batch_persist<FORM, MAXCPT, CW>), dim3(g), dim3(persist_block<CW>()), pad, s, a);
}
template <int FORM>
void launch_batch_persist_f(const StepDesc* dd, const uint8_t* gates, uint32_t* counters, int nd, uint32_t nitems,
                            int parity, int backwards, int flags, int max_cpt, uint32_t cap, hipStream_t s) {
  constexpr bool CW = persist_claim_wave<FORM>();
  if (max_cpt >= 4)
    launch_batch_persist_t<FORM, 4, CW>(dd, gates, counters, nd, nitems, parity, backwards, flags, cap, s);
  else if (max_cpt == 2)
    launch_batch_persist_t<FORM, 2, CW>(dd, gates, counters, nd, nitems, parity, backwards, flags, cap, s);
  else
    launch_batch_persist_t<FORM, 1, CW>(dd, gates, counters, nd, nitems, parity, backwards, flags, cap, s);
}

#define WT_BATCH_PERSIST_DECL(F)                                                                            \
  void launch_batch_persist_f<F>(const StepDesc* dd, const uint8_t* gates, uint32_t* counters, int nd,          \
                                 uint32_t nitems, int parity, int backwards, int flags, int max_cpt, uint32_t cap, \
                                 hipStream_t s)
#ifndef WT_STEP_PERSIST_UNIT
extern template WT_BATCH_PERSIST_DECL(WT_FP32);
extern template WT_BATCH_PERSIST_DECL(WT_FP16);
extern template WT_BATCH_PERSIST_DECL(WT_BF16);
extern template WT_BATCH_PERSIST_DECL(WT_FP64);
#endif

// One workgroup per item, or the persistent form (with at most cap workgroups
// when cap is positive).
template <int FORM>
void launch_batch_f(const StepDesc* dd, const uint8_t* gates, uint32_t* counters, int nd, uint32_t nitems,
                    int parity, int backwards, int flags, int max_cpt, int persist, uint32_t cap, hipStream_t s) {
  if (persist) {
    launch_batch_persist_f<FORM>(dd, gates, counters, nd, nitems, parity, backwards, flags, max_cpt, cap, s);
    return;
  }
  const size_t pad = batch_pad_lds<FORM>();
  if (max_cpt >= 4)
    hipLaunchKernelGGL((k_step_batch<FORM, 4>), dim3(nitems), dim3(BLOCK), pad, s, dd, gates, nd, nitems, parity,
                       backwards, flags);
  else if (max_cpt == 2)
    hipLaunchKernelGGL((k_step_batch<FORM, 2>), dim3(nitems), dim3(BLOCK), pad, s, dd, gates, nd, nitems, parity,
                       backwards, flags);
  else
    hipLaunchKernelGGL((k_step_batch<FORM, 1>), dim3(nitems), dim3(BLOCK), pad, s, dd, gates, nd, nitems, parity,
                       backwards, flags);
}

// The FP32 single-tunnel step kernels are instantiated in their own translation
// unit, wt_step_fp32.hip (this file with WT_STEP_FP32_UNIT defined), which the
// build compiles with the compiler's max-ILP scheduling: measured 1.7 % faster
// for the FP32 steps at 33.5M cells (5 waves per SIMD instead of 7), and 8 %
// slower for FP16, whose kernels keep the default scheduling here. The batched
// FP32 step stays in this unit: with max-ILP it takes 99 registers (4 waves per
// SIMD), with the default scheduling 80 (6 waves), 4.3 % faster on 48 tunnels.
#ifdef WT_STEP_FP32_UNIT
template void launch_step_f<WT_FP32>(const StepDesc& d, int parity, Shape sh, int flags, hipStream_t s);
template void launch_step_adv_f<WT_FP32>(const StepDesc& d, const AdvectArgs& a, Shape sh, int flags,
                                        hipStream_t s);

}  // namespace wt
#elif defined(WT_STEP_PERSIST_UNIT)
template WT_BATCH_PERSIST_DECL(WT_FP32);
template WT_BATCH_PERSIST_DECL(WT_FP16);
template WT_BATCH_PERSIST_DECL(WT_BF16);
template WT_BATCH_PERSIST_DECL(WT_FP64);

}  // namespace wt
#else
extern template void launch_step_f<WT_FP32>(const StepDesc& d, int parity, Shape sh, int flags, hipStream_t s);
extern template void launch_step_adv_f<WT_FP32>(const StepDesc& d, const AdvectArgs& a, Shape sh, int flags,
                                               hipStream_t s);

static int64_t nslots_of(int64_t n) { return (n + SLOT - 1) / SLOT; }
static int64_t nchunks_of(int64_t n) { return (nslots_of(n) + CHUNK - 1) / CHUNK; }

StepDesc make_desc(const wt_grid* g, int cur, const wt_params* p, int flags) {
  StepDesc d{};
  d.fin = g->f[cur & 1];
  d.fout = g->f[(cur & 1) ^ 1];
  d.cls = g->cls;
  d.mac = g->mac;
  d.shell_in = g->shell[cur & 1];
  d.shell_out = g->shell[(cur & 1) ^ 1];
  d.partials = g->partials;
  const Div dx = make_div((uint32_t)g->nx), dy = make_div((uint32_t)g->ny);
  d.mx = dx.m; d.sx = dx.s; d.my = dy.m; d.sy = dy.s;
  d.nx = g->nx; d.ny = g->ny; d.nz = g->nz;
  d.n = (uint32_t)((int64_t)g->nx * g->ny * g->nz);
  if (p) { d.uin = p->uin; d.tau = p->tau; d.smag = p->smag; }
  d.tile_begin = 0;
  d.ntiles = 0;
  d.flags = flags;
  // raw-buffer addressing needs the population bytes plus the scalar-offset bias in 32 bits
  const int64_t eb = form_elem_bytes(g->form);
  const int64_t span = (int64_t)Q * d.n * eb + (int64_t)Q * ((int64_t)g->nx * g->ny + g->nx) * eb;
  d.global_addr = (eb > 4 || span >= ((int64_t)1 << 32)) ? 1 : 0;
  return d;
}

int check_grid(const wt_grid* g, bool need_state) {
  if (!g) return fail(-1, "null grid");
  if (g->nx < 3 || g->ny < 3 || g->nz < 3)
    return fail(-1, "grid dimensions must be at least 3 (got %d x %d x %d)", g->nx, g->ny, g->nz);
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  if (n >= (int64_t)1 << 31) return fail(-1, "grid of %lld cells exceeds 2^31 - 1", (long long)n);
  if ((int64_t)Q * n >= ((int64_t)1 << 32))
    return fail(-1, "grid of %lld cells: 19*n population elements exceed 2^32", (long long)n);
  if (g->form < WT_FP32 || g->form > WT_FP64) return fail(-1, "unknown storage form %d", g->form);
  if (need_state && (!g->f[0] || !g->f[1] || !g->cls || !g->mac || !g->shell[0] || !g->shell[1]))
    return fail(-1, "grid has a null device buffer");
  // the kernels read the face snapshot four reals at a time, store the macroscopic
  // field as one vector per cell and the populations and class bytes through
  // dword-addressed buffers: every state buffer 16-byte aligned (the FP64
  // snapshot 32), as any device allocation is
  if (need_state) {
    const uintptr_t a = (uintptr_t)g->f[0] | (uintptr_t)g->f[1] | (uintptr_t)g->cls | (uintptr_t)g->mac;
    const uintptr_t sh = (uintptr_t)g->shell[0] | (uintptr_t)g->shell[1];
    if ((a & 15) || (sh & (g->form == WT_FP64 ? 31 : 15)) || ((uintptr_t)g->partials & 7) || ((uintptr_t)g->sums & 7))
      return fail(-1, "grid buffers must be 16-byte aligned (the FP64 face snapshot 32-byte)");
  }
  return 0;
}

// ---- scratch of the allocation-free step ------------------------------------------
// wt_step_noalloc keeps the FP64 reduction slots of the force/mass epilogue
// (4 doubles per 256-cell slot, then 4 per 1024-slot chunk) in this array
// instead of the grid's `partials` buffer: 8.4 MB on every device the library
// runs on, held until the process ends, enough for tunnels of up to 2^26 cells.
// Every slot and chunk sum a call reads is written earlier in the same call, so
// nothing is cleared; calls that overlap on one device share it and race.
constexpr int64_t NOALLOC_MAX_CELLS = (int64_t)1 << 26;
constexpr int64_t NOALLOC_SLOTS = NOALLOC_MAX_CELLS / SLOT;
constexpr int64_t NOALLOC_CHUNKS = NOALLOC_SLOTS / CHUNK;
static __device__ __align__(256) double scratch[4 * (NOALLOC_SLOTS + NOALLOC_CHUNKS)];

// Device address of `scratch` on the current device, looked up once per device
// with that device current (each device holds its own copy at its own address).
// Host code uses this pointer, never the array's name, for the array's memory.
static int scratch_ptr(double** out) {
  static std::mutex lock;
  static double* ptrs[64] = {nullptr};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(-2, "hipGetDevice failed: %s", hipGetErrorString(hipGetLastError()));
  if (dev < 0 || dev >= 64) return fail(-1, "device %d beyond the 64 this library tracks", dev);
  std::lock_guard<std::mutex> g(lock);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(-2, "hipGetSymbolAddress of the step scratch failed: %s", hipGetErrorString(e));
    ptrs[dev] = (double*)p;
  }
  *out = ptrs[dev];
  return 0;
}

static int reduce_one(const wt_grid* g, hipStream_t s) {
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  if (!g->partials || !g->sums) return fail(-1, "grid has no reduction buffers");
  RedDesc r{g->partials, g->sums, g->sums, (uint32_t)nslots_of(n)};
  hipLaunchKernelGGL(k_reduce1_one, dim3((unsigned)nchunks_of(n)), dim3(BLOCK), 0, s, r);
  hipLaunchKernelGGL(k_reduce2_one, dim3(1), dim3(BLOCK), 0, s, r);
  return check_launch("reduce");
}


// Launch shape of a tunnel's step kernels from its grid (cpt 0 means 1).
static int grid_shape(const wt_grid* g, Shape* sh) {
  const int cpt = g->cpt == 0 ? 1 : g->cpt;
  if (cpt != 1 && cpt != 2 && cpt != 4) return fail(-1, "cells per thread must be 1, 2 or 4 (got %d)", g->cpt);
  if (g->tile_order & ~7) return fail(-1, "unknown tile order bits 0x%x", g->tile_order);
  sh->cpt = cpt;
  sh->remap = g->tile_order;
  sh->solo = 0;
  return 0;
}

// Cache policy of a single-tunnel launch (run_tile's POL) by the bytes of its two
// population buffers. Streaming (1) pays only far above the 256 MB Infinity Cache:
// from 2 GiB (FP32 ~14M cells, FP16 ~28M); against policy 0 it measured 33.5M cells
// 0.9-1.05 % (FP32) / 0.7-1.2 % (FP16) faster, 11.8M even / 1.5 % slower, 3.9M
// 5.8 % / 8.4 % slower, 0.5M 2 % slower. All-cached (2) pays up to 224 MiB of
// populations: faster than policy 0 by 8.8 % (FP32) / 6.5 % (FP16) at 0.5M cells,
// 7.9 % (FP32) at 1.0M, 8.5 % (FP32, 196 MiB) at 1.35M, 12.5 % (FP16, 128 MiB) at
// 1.77M; slower by 3.1 % (FP32, 257 MiB) at 1.77M and 4-16 % at 3.9M-11.8M cells.
static int solo_policy(const wt_grid* g) {
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  const int64_t bytes = 2 * (int64_t)Q * n * form_elem_bytes(g->form);
  if (bytes >= ((int64_t)1 << 31)) return 1;
  return bytes <= ((int64_t)224 << 20) ? 2 : 0;
}

static void launch_step(int form, const StepDesc& d, int parity, Shape sh, int flags, hipStream_t s) {
  switch (form) {
    case WT_FP32: launch_step_f<WT_FP32>(d, parity, sh, flags, s); break;
    case WT_FP16: launch_step_f<WT_FP16>(d, parity, sh, flags, s); break;
    case WT_BF16: launch_step_f<WT_BF16>(d, parity, sh, flags, s); break;
    case WT_FP64: launch_step_f<WT_FP64>(d, parity, sh, flags, s); break;
  }
}

}  // namespace wt

using namespace wt;

extern "C" {

int wt_cache_policy(int nx, int ny, int nz, int form) {
  wt_grid g{};
  g.nx = nx; g.ny = ny; g.nz = nz; g.form = form;
  if (int e = check_grid(&g, false)) return e;
  return solo_policy(&g);
}

int wt_buffer_sizes(int nx, int ny, int nz, int form, int64_t sizes[6]) {
  if (nx < 3 || ny < 3 || nz < 3) return fail(-1, "grid dimensions must be at least 3");
  if (form < WT_FP32 || form > WT_FP64) return fail(-1, "unknown storage form %d", form);
  const int64_t n = (int64_t)nx * ny * nz;
  sizes[0] = (int64_t)Q * n * form_elem_bytes(form);
  sizes[1] = n;
  sizes[2] = mac_elem_bytes(form) * n;
  sizes[3] = 4 * shell_entries(nx, ny, nz) * form_real_bytes(form);
  sizes[4] = 32 * (nslots_of(n) + nchunks_of(n));
  sizes[5] = 32;
  return 0;
}

int64_t wt_state_bytes(int nx, int ny, int nz, int form) {
  int64_t s[6];
  if (wt_buffer_sizes(nx, ny, nz, form, s) != 0) return -1;
  return 2 * s[0] + s[1] + s[2] + 2 * s[3] + s[4] + s[5];
}

int64_t wt_tile_count(int nx, int ny, int nz) { return nslots_of((int64_t)nx * ny * nz); }

int wt_init(const wt_grid* g, float uin, void* stream) {
  if (int e = check_grid(g, true)) return e;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = (uint32_t)((int64_t)g->nx * g->ny * g->nz);
  const dim3 grid((n + BLOCK - 1) / BLOCK);
  void* mac = g->mac;
  const uint32_t ent = (uint32_t)shell_entries(g->nx, g->ny, g->nz);
  const dim3 sgrid((ent + BLOCK - 1) / BLOCK);
  switch (g->form) {
    case WT_FP32:
      hipLaunchKernelGGL(k_init<WT_FP32>, grid, dim3(BLOCK), 0, s, (float*)g->f[0], (float*)g->f[1], mac, n, (uint32_t)g->nx, uin);
      break;
    case WT_FP16:
      hipLaunchKernelGGL(k_init<WT_FP16>, grid, dim3(BLOCK), 0, s, (uint16_t*)g->f[0], (uint16_t*)g->f[1], mac, n, (uint32_t)g->nx, uin);
      break;
    case WT_BF16:
      hipLaunchKernelGGL(k_init<WT_BF16>, grid, dim3(BLOCK), 0, s, (uint16_t*)g->f[0], (uint16_t*)g->f[1], mac, n, (uint32_t)g->nx, uin);
      break;
    case WT_FP64:
      hipLaunchKernelGGL(k_init<WT_FP64>, grid, dim3(BLOCK), 0, s, (double*)g->f[0], (double*)g->f[1], mac, n, (uint32_t)g->nx, uin);
      break;
  }
  if (g->form == WT_FP64)
    hipLaunchKernelGGL(k_init_shell<double>, sgrid, dim3(BLOCK), 0, s, (double*)g->shell[0], (double*)g->shell[1], ent, uin);
  else
    hipLaunchKernelGGL(k_init_shell<float>, sgrid, dim3(BLOCK), 0, s, (float*)g->shell[0], (float*)g->shell[1], ent, uin);
  return check_launch("wt_init");
}

int wt_set_solid(const wt_grid* g, const uint8_t* solid_dev, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!solid_dev || !g->cls) return fail(-1, "null solid mask or class buffer");
  const uint32_t n = (uint32_t)((int64_t)g->nx * g->ny * g->nz);
  const Div dx = make_div((uint32_t)g->nx), dy = make_div((uint32_t)g->ny);
  hipLaunchKernelGGL(k_build_cls, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream,
                     solid_dev, g->cls, g->nx, g->ny, g->nz, n, dx.m, dx.s, dy.m, dy.s);
  return check_launch("wt_set_solid");
}

int wt_step(const wt_grid* g, int cur, int nsteps, const wt_params* p, int flags, void* stream) {
  if (int e = check_grid(g, true)) return e;
  if (!p) return fail(-1, "null params");
  if (nsteps < 0) return fail(-1, "negative step count %d", nsteps);
  if (flags & ~(WT_FLAG_FORCES | WT_FLAG_MASS | WT_FLAG_NO_MAC)) return fail(-1, "unknown flag bits 0x%x", flags);
  if ((flags & (WT_FLAG_FORCES | WT_FLAG_MASS)) && (!g->partials || !g->sums))
    return fail(-1, "reduction flags need partials and sums buffers");
  Shape sh;
  if (int e = grid_shape(g, &sh)) return e;
  hipStream_t s = (hipStream_t)stream;
  const int solo = solo_policy(g);
  StepDesc d = make_desc(g, cur, p, 0);
  if (sh.remap & 4) d.global_addr = 1;  // shape bit 2: 64-bit global addressing at any size
  for (int k = 0; k < nsteps; k++) {
    const int f = ((k == nsteps - 1) ? (flags & (WT_FLAG_FORCES | WT_FLAG_MASS)) : 0) | (flags & WT_FLAG_NO_MAC);
    // with shape bit 1 the sweep direction alternates with the buffer parity
    Shape shk = sh;
    shk.remap = (sh.remap & 1) | (((sh.remap & 2) && ((cur ^ k) & 1)) ? 2 : 0);
    shk.solo = solo;
    launch_step(g->form, d, k & 1, shk, f, s);
  }
  if (nsteps > 0 && (flags & (WT_FLAG_FORCES | WT_FLAG_MASS))) {
    if (int e = check_launch("wt_step")) return e;
    return reduce_one(g, s);
  }
  return check_launch("wt_step");
}

int wt_step_noalloc(const wt_grid* g, int cur, int nsteps, const wt_params* p, int flags, void* stream) {
  if (int e = check_grid(g, false)) return e;
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  if (n > NOALLOC_MAX_CELLS)
    return fail(-1, "wt_step_noalloc: %lld cells exceed its scratch's maximum of %lld; use wt_step",
                (long long)n, (long long)NOALLOC_MAX_CELLS);
  double* sp = nullptr;
  if (int e = scratch_ptr(&sp)) return e;
  wt_grid h = *g;
  h.partials = sp;       // the reduction slots live in the static scratch
  return wt_step(&h, cur, nsteps, p, flags, stream);
}

int wt_step_advect(const wt_grid* g, int cur, const wt_params* p, int flags, const void* mac_prev, float* pos,
                   float* spd, const uint32_t* ids, int64_t num_p, float dt, float uin, float max_age,
                   uint32_t seed, const wt_seed_box* box, void* stream) {
  if (int e = check_grid(g, true)) return e;
  if (!p) return fail(-1, "null params");
  if ((flags & (WT_FLAG_FORCES | WT_FLAG_MASS)) && (!g->partials || !g->sums))
    return fail(-1, "reduction flags need partials and sums buffers");
  if (flags & ~(WT_FLAG_FORCES | WT_FLAG_MASS))
    return fail(-1, "wt_step_advect takes WT_FLAG_FORCES and WT_FLAG_MASS only (flags 0x%x; WT_FLAG_NO_MAC "
                    "would leave unwritten the field the next step's advection reads)", flags);
  if (num_p < 0 || num_p >= ((int64_t)1 << 31)) return fail(-1, "bad particle count %lld", (long long)num_p);
  if (num_p > 0 && (!pos || !spd || !box || !mac_prev)) return fail(-1, "null particle buffer, field or seed box");
  if (num_p > 0 && mac_prev == (const void*)g->mac)
    return fail(-1, "the advected field must not be the grid's mac buffer (the step writes it)");
  Shape sh;
  if (int e = grid_shape(g, &sh)) return e;
  hipStream_t s = (hipStream_t)stream;
  StepDesc d = make_desc(g, cur, p, 0);
  if (sh.remap & 4) d.global_addr = 1;
  sh.remap = (sh.remap & 1) | (((sh.remap & 2) && (cur & 1)) ? 2 : 0);   // as step k = 0 of wt_step
  sh.solo = solo_policy(g);
  const wt_seed_box none{0.f, 0.f, 0.f, 0.f, 0.f};
  const AdvectArgs a{MacView{mac_prev, mac_is_half(g->form)}, g->cls, nullptr, ids, reinterpret_cast<float4*>(pos),
                     spd, (uint32_t)num_p, g->nx, g->ny, g->nz, dt, uin, max_age, seed, box ? *box : none};
  const int f = flags & (WT_FLAG_FORCES | WT_FLAG_MASS);
  switch (g->form) {
    case WT_FP32: launch_step_adv_f<WT_FP32>(d, a, sh, f, s); break;
    case WT_FP16: launch_step_adv_f<WT_FP16>(d, a, sh, f, s); break;
    case WT_BF16: launch_step_adv_f<WT_BF16>(d, a, sh, f, s); break;
    case WT_FP64: launch_step_adv_f<WT_FP64>(d, a, sh, f, s); break;
  }
  if (int e = check_launch("wt_step_advect")) return e;
  return f ? reduce_one(g, s) : 0;
}

int wt_forces(const wt_grid* g, int cur, void* stream) {
  if (int e = check_grid(g, true)) return e;
  if (!g->partials || !g->sums) return fail(-1, "grid has no reduction buffers");
  hipStream_t s = (hipStream_t)stream;
  StepDesc d = make_desc(g, cur, nullptr, 0);
  const dim3 grid((d.n + BLOCK - 1) / BLOCK);
  switch (g->form) {
    case WT_FP32: hipLaunchKernelGGL(k_force<WT_FP32>, grid, dim3(BLOCK), 0, s, d); break;
    case WT_FP16: hipLaunchKernelGGL(k_force<WT_FP16>, grid, dim3(BLOCK), 0, s, d); break;
    case WT_BF16: hipLaunchKernelGGL(k_force<WT_BF16>, grid, dim3(BLOCK), 0, s, d); break;
    case WT_FP64: hipLaunchKernelGGL(k_force<WT_FP64>, grid, dim3(BLOCK), 0, s, d); break;
  }
  if (int e = check_launch("wt_forces")) return e;
  return reduce_one(g, s);
}

int wt_read_sums(const wt_grid* g, double out[4], void* stream) {
  if (!g || !g->sums || !out) return fail(-1, "null grid, sums or output");
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(out, g->sums, 4 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
    return fail(-2, "hipMemcpyAsync of sums failed: %s", hipGetErrorString(hipGetLastError()));
  if (hipStreamSynchronize(s) != hipSuccess)
    return fail(-2, "stream synchronize failed: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}

int wt_link_classes(const wt_grid* g, uint64_t* out_dev, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!out_dev || !g->cls) return fail(-1, "null output or class buffer");
  const uint32_t n = (uint32_t)((int64_t)g->nx * g->ny * g->nz);
  const Div dx = make_div((uint32_t)g->nx), dy = make_div((uint32_t)g->ny);
  hipLaunchKernelGGL(k_link_classes, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream,
                     g->cls, out_dev, g->nx, g->ny, g->nz, n, dx.m, dx.s, dy.m, dy.s);
  return check_launch("wt_link_classes");
}

// ---- batched tunnels ---------------------------------------------------------------------
// Workspace layout: [StepDesc x ntunnels][RedDesc x ntunnels][gate bytes, padded
// to 8][up to 128 zero bytes][the persistent kernel's claim counters, one per
// 128-byte line].
static int64_t batch_counter_offset(int ntunnels) {
  const int64_t end = wt_batch_gate_offset(ntunnels) + (((int64_t)ntunnels + 7) & ~(int64_t)7);
  return (end + 127) & ~(int64_t)127;
}

int64_t wt_batch_workspace_bytes(int ntunnels) {
  if (ntunnels <= 0) {
    fail(-1, "ntunnels must be positive");
    return -1;
  }
  return batch_counter_offset(ntunnels) + (int64_t)BATCH_COUNTERS * BATCH_COUNTER_STRIDE * 4;
}

int64_t wt_batch_gate_offset(int ntunnels) {
  if (ntunnels <= 0) {
    fail(-1, "ntunnels must be positive");
    return -1;
  }
  return (int64_t)ntunnels * (int64_t)(sizeof(StepDesc) + sizeof(RedDesc));
}

int wt_batch_encode(const wt_grid* grids, const wt_params* params, const int* cur, int ntunnels,
                    double* sums_dev, void* host_image, wt_batch_info* info) {
  if (!grids || !params || !cur || !host_image || !info || ntunnels <= 0)
    return fail(-1, "null argument or empty batch");
  const int form = grids[0].form;
  if (form < WT_FP32 || form > WT_FP64) return fail(-1, "unknown storage form %d", form);
  StepDesc* ds = (StepDesc*)host_image;
  RedDesc* rs = (RedDesc*)((char*)host_image + sizeof(StepDesc) * ntunnels);
  uint8_t* gates = (uint8_t*)host_image + wt_batch_gate_offset(ntunnels);
  uint64_t tile = 0;
  int64_t max_chunks = 0;
  int max_cpt = 1;
  for (int k = 0; k < ntunnels; k++) {
    if (int e = check_grid(&grids[k], true)) return e;
    if (grids[k].form != form)
      return fail(-1, "tunnel %d has storage form %d, batch uses %d", k, grids[k].form, form);
    if (!grids[k].partials || !grids[k].sums) return fail(-1, "tunnel %d has no reduction buffers", k);
    Shape sh;
    if (int e = grid_shape(&grids[k], &sh)) return e;
    StepDesc d = make_desc(&grids[k], 0, &params[k], 0);
    if (sh.remap & 4) d.global_addr = 1;
    d.cpt = sh.cpt;
    max_cpt = std::max(max_cpt, sh.cpt);
    d.tile_begin = (uint32_t)tile;
    d.ntiles = (d.n + BLOCK * sh.cpt - 1) / (BLOCK * sh.cpt);
    d.flags = cur[k] & 1;
    tile += d.ntiles;
    if (tile >= (1ull << 31)) return fail(-1, "batch has too many tiles");
    ds[k] = d;
    if (sums_dev)
      rs[k] = RedDesc{grids[k].partials, sums_dev + 4 * (size_t)k, sums_dev + 4 * ((size_t)ntunnels + k),
                      (uint32_t)nslots_of(d.n)};
    else
      rs[k] = RedDesc{grids[k].partials, grids[k].sums, grids[k].sums, (uint32_t)nslots_of(d.n)};
    max_chunks = std::max<int64_t>(max_chunks, nchunks_of(d.n));
    gates[k] = 1;  // every tunnel stores its macroscopic field until the caller says otherwise
  }
  for (int64_t k = ntunnels; k < ((ntunnels + 7) & ~7); k++) gates[k] = 0;
  // the persistent kernel's claim counters start at 0 (and it leaves them at 0)
  std::memset((char*)host_image + wt_batch_gate_offset(ntunnels) + ((ntunnels + 7) & ~7), 0,
              (size_t)(wt_batch_workspace_bytes(ntunnels) - wt_batch_gate_offset(ntunnels) - ((ntunnels + 7) & ~7)));
  info->total_tiles = (int64_t)tile;
  info->max_chunks = (int32_t)max_chunks;
  info->form = form;
  info->ntunnels = ntunnels;
  info->tile_order = WT_BATCH_BACKWARDS;
  info->max_cpt = max_cpt;
  info->persist_blocks = 0;
  return 0;
}

int wt_batch_prepare(const wt_grid* grids, const wt_params* params, const int* cur, int ntunnels,
                     double* sums_dev, void* workspace, wt_batch_info* info, void* stream) {
  if (!workspace || ntunnels <= 0) return fail(-1, "null workspace or empty batch");
  if ((uintptr_t)workspace & 15) return fail(-1, "batch workspace is not 16-byte aligned");
  std::vector<char> image((size_t)wt_batch_workspace_bytes(ntunnels));
  if (int e = wt_batch_encode(grids, params, cur, ntunnels, sums_dev, image.data(), info)) return e;
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(workspace, image.data(), image.size(), hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(-2, "descriptor upload failed: %s", hipGetErrorString(hipGetLastError()));
  if (hipStreamSynchronize(s) != hipSuccess)
    return fail(-2, "stream synchronize failed: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}

// One step of every tunnel prepared in `workspace`: with parity p each tunnel
// k advances from its buffer cur[k] ^ p. Kernel launches only, so a caller
// may capture it in a graph (two calls, p = 0 and p = 1, per replay).
int wt_batch_step(void* workspace, const wt_batch_info* info, int parity, int flags, void* stream) {
  if (!workspace || !info || info->ntunnels <= 0 || info->total_tiles <= 0) return fail(-1, "empty batch");
  if ((uintptr_t)workspace & 15) return fail(-1, "batch workspace is not 16-byte aligned");
  const int form = info->form;
  if (form < WT_FP32 || form > WT_FP64) return fail(-1, "unknown storage form %d", form);
  if (info->tile_order & ~(WT_BATCH_BACKWARDS | WT_BATCH_PERSISTENT))
    return fail(-1, "unknown batch tile order bits 0x%x", info->tile_order);
  if (info->max_cpt != 1 && info->max_cpt != 2 && info->max_cpt != 4)
    return fail(-1, "batch info has cells per thread %d (not set by wt_batch_encode?)", info->max_cpt);
  if (flags & ~(WT_FLAG_FORCES | WT_FLAG_MASS | WT_FLAG_NO_MAC)) return fail(-1, "unknown flag bits 0x%x", flags);
  hipStream_t s = (hipStream_t)stream;
  const int nt = info->ntunnels;
  const StepDesc* dd = (const StepDesc*)workspace;
  const RedDesc* rd = (const RedDesc*)((const char*)workspace + sizeof(StepDesc) * nt);
  const uint8_t* gates = (const uint8_t*)workspace + wt_batch_gate_offset(nt);
  uint32_t* ctrs = (uint32_t*)((char*)workspace + batch_counter_offset(nt));
  const int f = flags & (WT_FLAG_FORCES | WT_FLAG_MASS | WT_FLAG_NO_MAC);
  const uint32_t nitems = (uint32_t)info->total_tiles;
  const int backwards = ((info->tile_order & WT_BATCH_BACKWARDS) && (parity & 1)) ? 1 : 0;
  if (info->persist_blocks < 0) return fail(-1, "negative persist_blocks %d", info->persist_blocks);
  const int persist = (info->tile_order & WT_BATCH_PERSISTENT) ? 1 : 0;
  const uint32_t cap = (uint32_t)info->persist_blocks;
  switch (form) {
    case WT_FP32: launch_batch_f<WT_FP32>(dd, gates, ctrs, nt, nitems, parity, backwards, f, info->max_cpt, persist, cap, s); break;
    case WT_FP16: launch_batch_f<WT_FP16>(dd, gates, ctrs, nt, nitems, parity, backwards, f, info->max_cpt, persist, cap, s); break;
    case WT_BF16: launch_batch_f<WT_BF16>(dd, gates, ctrs, nt, nitems, parity, backwards, f, info->max_cpt, persist, cap, s); break;
    case WT_FP64: launch_batch_f<WT_FP64>(dd, gates, ctrs, nt, nitems, parity, backwards, f, info->max_cpt, persist, cap, s); break;
  }
  if (f & (WT_FLAG_FORCES | WT_FLAG_MASS)) {
    hipLaunchKernelGGL(k_reduce1, dim3((unsigned)info->max_chunks, (unsigned)nt), dim3(BLOCK), 0, s, rd);
    hipLaunchKernelGGL(k_reduce2, dim3((unsigned)nt), dim3(BLOCK), 0, s, rd, parity & 1);
  }
  return check_launch("wt_batch_step");
}

}  // extern "C"

#endif  // WT_STEP_FP32_UNIT
