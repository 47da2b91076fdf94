// Device-side building blocks of the D3Q19 solver for gfx950: lattice tables,
// population storage forms, grid descriptors and integer helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wt.h"

namespace wt {

constexpr int Q = 19;

// D3Q19 velocity set in the reference's order (lbm.py:15-17): 0 rest,
// 1-6 axis +-x, +-y, +-z, 7-18 diagonals in opposite pairs.
constexpr int cx(int i) {
  constexpr int t[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[i];
}
constexpr int cy(int i) {
  constexpr int t[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[i];
}
constexpr int cz(int i) {
  constexpr int t[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  return t[i];
}
// opposite direction, lbm.py:19
constexpr int opp(int i) {
  constexpr int t[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
  return t[i];
}
// lattice weight in double, lbm.py:18
// Population layout: row-blocked SoA. Each x-row r = y + ny*z of nx cells
// stores its 19 planes of nx values one after another: population i of cell
// idx = x + nx*r is element 19*nx*r + nx*ppos(i) + x. A wave's consecutive cells read
// or write one contiguous run per plane, and a row's 19 runs lie together.
// Interior cells never pull across their own row in x and y/z neighbours are
// whole rows away, so population i of the upstream neighbour idx - c_i sits at
// the fixed offset pshift(i) from the cell's own plane-0 element prow(idx, r).
__host__ __device__ __forceinline__ uint32_t prow(uint32_t idx, uint32_t row, uint32_t nx) {
  return idx + (uint32_t)(Q - 1) * nx * row;
}
// Position of direction i's plane inside a row block. Planes are grouped by the
// row that reads them (equal (c_y, c_z)): {0, 1, 2}, {3, 7, 10}, {4, 8, 9},
// {5, 11, 14}, {6, 12, 13}, 15, 16, 17, 18, so the pulls of one wave from one
// upstream row come from one contiguous stretch of that row's block.
__host__ __device__ constexpr int ppos(int i) {
  constexpr int P[Q] = {0, 1, 2, 3, 6, 9, 12, 4, 7, 8, 5, 10, 13, 14, 11, 15, 16, 17, 18};
  return P[i];
}
// pshift in 32-bit wrap-around arithmetic: prow + pshift is the exact element
// index (< 19 n < 2^32) even where the signed offset is negative.
__host__ __device__ __forceinline__ uint32_t pshift(int i, uint32_t nx, uint32_t nxy) {
  return (uint32_t)ppos(i) * nx - ((uint32_t)cx(i) + (uint32_t)(Q * cy(i)) * nx + (uint32_t)(Q * cz(i)) * nxy);
}

constexpr double wd(int i) { return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

// class byte written by wt_set_solid
constexpr uint8_t CLS_CODE = 3;   // 0 fluid, 1 static solid, 2 belt, 3 other solid
constexpr uint8_t CLS_NEAR = 128; // interior fluid cell with a non-fluid neighbour
constexpr uint8_t CLS_FORCE = 64; // interior fluid cell with a static-solid neighbour

// ---- storage forms ----------------------------------------------------------
// load/store: the population f_i of a stored element (weight added back for the
// deviation forms). pop/put: the value the collision works on -- f_i for FP32
// and FP64, the deviation g_i = f_i - w_i for FP16 and BF16, whose collision
// runs on deviations throughout (dev = true). excess: f_i - w_i of a stored
// element, the momentum-exchange term of lbm.py:189.
template <int FORM> struct Form;

template <> struct Form<WT_FP32> {
  typedef float elem;
  typedef float real;
  static constexpr bool dev = false;
  __device__ static float load(elem e, int) { return e; }
  __device__ static elem store(float f, int) { return f; }
  __device__ static float pop(elem e) { return e; }
  __device__ static elem put(float v) { return v; }
  __device__ static float excess(elem e, int i) { return e - (float)wd(i); }
};

__device__ inline float h2f(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }
__device__ inline uint16_t f2h(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// Round to nearest even, NaN kept quiet: gfx950's v_cvt_pk_bf16_f32, one
// instruction where the bit arithmetic of the same rounding takes five or six
// (x += 0x7fff + lsb, x >> 16; NaN: (x >> 16) | 0x40, the oracle's form).
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

template <> struct Form<WT_FP16> {
  typedef uint16_t elem;
  typedef float real;
  static constexpr bool dev = true;
  __device__ static float load(elem e, int i) { return (float)wd(i) + h2f(e); }
  __device__ static elem store(float f, int i) { return f2h(f - (float)wd(i)); }
  __device__ static float pop(elem e) { return h2f(e); }
  __device__ static elem put(float v) { return f2h(v); }
  __device__ static float excess(elem e, int) { return h2f(e); }
};

template <> struct Form<WT_BF16> {
  typedef uint16_t elem;
  typedef float real;
  static constexpr bool dev = true;
  __device__ static float load(elem e, int i) { return (float)wd(i) + bf2f(e); }
  __device__ static elem store(float f, int i) { return f2bf(f - (float)wd(i)); }
  __device__ static float pop(elem e) { return bf2f(e); }
  __device__ static elem put(float v) { return f2bf(v); }
  __device__ static float excess(elem e, int) { return bf2f(e); }
};

template <> struct Form<WT_FP64> {
  typedef double elem;
  typedef double real;
  static constexpr bool dev = false;
  __device__ static double load(elem e, int) { return e; }
  __device__ static elem store(double f, int) { return f; }
  __device__ static double pop(elem e) { return e; }
  __device__ static elem put(double v) { return v; }
  __device__ static double excess(elem e, int i) { return e - wd(i); }
};

// ---- the macroscopic field ------------------------------------------------------
// FP32 and FP64 tunnels store (ux, uy, uz, rho) as float4 per cell (lbm.py:406);
// FP16 and BF16 tunnels store it as 4 binary16 values (ux, uy, uz, rho - 1),
// each rounded to nearest even from FP32, 8 B per cell instead of 16. Readers
// rebuild float4 (ux, uy, uz, 1 + (rho - 1)) in FP32. The step never reads it
// (face cells read the FP32 face-layer snapshot), so its precision changes no
// population; particles, slices and Cp read it.
inline bool mac_is_half(int form) { return form == WT_FP16 || form == WT_BF16; }
inline int mac_elem_bytes(int form) { return mac_is_half(form) ? 8 : 16; }

struct MacView {
  const void* p;
  int half;
  __device__ __forceinline__ float4 operator[](uint32_t idx) const {
    if (half) {
      const uint2 v = reinterpret_cast<const uint2*>(p)[idx];
      return make_float4(h2f((uint16_t)(v.x & 0xffffu)), h2f((uint16_t)(v.x >> 16)), h2f((uint16_t)(v.y & 0xffffu)),
                         1.0f + h2f((uint16_t)(v.y >> 16)));
    }
    return reinterpret_cast<const float4*>(p)[idx];
  }
};

inline int form_elem_bytes(int form) { return form == WT_FP32 ? 4 : (form == WT_FP64 ? 8 : 2); }
inline int form_real_bytes(int form) { return form == WT_FP64 ? 8 : 4; }

// ---- exact unsigned division by a runtime constant ---------------------------
// q = floor(x / d) for every x < 2^31, with m = ceil(2^(31+s) / d) < 2^32 and
// s = ceil(log2 d) (round-up reciprocal method for 31-bit dividends).
struct Div {
  uint64_t m;
  uint32_t s;
};
inline Div make_div(uint32_t d) {
  uint32_t s = 0;
  while ((1ull << s) < d) s++;
  uint64_t num = 1ull << (31 + s);
  uint64_t m = (num + d - 1) / d;
  return Div{m, 31 + s};
}
__device__ __forceinline__ uint32_t divu(uint32_t x, uint64_t m, uint32_t sh) {
  return (uint32_t)(((uint64_t)x * m) >> sh);
}

// ---- per-tunnel descriptor used by the step kernels ----------------------------
struct StepDesc {
  const void* fin;
  void* fout;
  const uint8_t* cls;
  void* mac;          // float4 or 4 x binary16 per cell by storage form (MacView)
  const void* shell_in;
  void* shell_out;
  double* partials;   // per-tile slots of 4 doubles, or null
  uint64_t mx, my;    // magic multipliers for / nx and / ny
  uint32_t sx, sy;
  int32_t nx, ny, nz;
  uint32_t n;
  float uin, tau, smag;
  uint32_t tile_begin; // first global tile of this tunnel (batched launches)
  uint32_t ntiles;
  int32_t flags;
  int32_t global_addr; // 1: populations beyond 32-bit buffer offsets, use 64-bit global addresses
  int32_t cpt;         // cells per thread of this tunnel's tiles (batched launches: 1, 2 or 4)
};

// Shell layout: six layers next to the faces, 4 reals per entry.
// xlo (x=1) [ny*nz], xhi (x=nx-2) [ny*nz], ylo (y=1) [nx*nz], yhi (y=ny-2) [nx*nz],
// zlo (z=1) [nx*ny], zhi (z=nz-2) [nx*ny].
inline int64_t shell_entries(int nx, int ny, int nz) {
  return 2 * ((int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny);
}

}  // namespace wt
