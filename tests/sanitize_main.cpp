// Host C++ under AddressSanitizer and UndefinedBehaviorSanitizer: the
// voxelizer of libwt (wind-tunnel_amd/csrc/wt_voxel.cpp) and the CPU oracle
// (oracle/wt_oracle.cpp), built together with this driver by g++ with
// -fsanitize=address,undefined (tests/test_sanitize_cpu.py). Any invalid
// access, overflow or undefined operation aborts with a sanitizer report.
//
//   sanitize_main <dir>
//     <dir>/index.txt: one line per voxelizer case,
//       name nx ny nz ground ntris count frontal planform
//     <dir>/<name>.tris (float32 triangles), <dir>/<name>.mask (uint8 mask)
// Prints one line per check and "sanitize ok" at the end.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "wt.h"

namespace wt {
char* error_buffer() {
  static thread_local char buf[512];
  return buf;
}
}  // namespace wt

extern "C" {
void oracle_set_threads(int n);
void oracle_init(int form, void* f, void* mac, int64_t n, float uin);
void oracle_step(int form, const void* fin, void* fout, const uint8_t* solid, const void* mac_prev, void* mac_out,
                 int nx, int ny, int nz, float uin, float tau, float smag, double* mass);
void oracle_force(int form, const void* f, const uint8_t* solid, int nx, int ny, int nz, double* out);
void oracle_link_classes(const uint8_t* solid, int nx, int ny, int nz, uint64_t* out);
void oracle_advect(float* pos, float* spd, const float* mac, const uint8_t* solid, int64_t num_p, int nx, int ny,
                   int nz, float dt, float uin, float max_age, uint32_t seed, float seed_x, float seed_cy,
                   float seed_cz, float seed_hy, float seed_hz);
void oracle_pack_particles(const float* pos, const float* spd, uint16_t* q16, uint8_t* q8, int64_t num_p, int nx,
                           int ny, int nz);
void oracle_slice(const float* mac, const uint8_t* solid, uint8_t* out, int nx, int ny, int nz, int axis, int index,
                  int field, float uin, float lref, float vmin, float vmax);
void oracle_cp(const float* mac, const uint8_t* solid, const float* verts, const float* norms, int8_t* out,
               int64_t nv, int nx, int ny, int nz, float uin, float cp_range);
}

static std::vector<char> slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

static int voxel_cases(const std::string& dir) {
  std::ifstream idx(dir + "/index.txt");
  std::string name;
  int nx, ny, nz, ground;
  long long ntris, count, frontal, planform;
  int ncase = 0;
  while (idx >> name >> nx >> ny >> nz >> ground >> ntris >> count >> frontal >> planform) {
    const std::vector<char> tris = slurp(dir + "/" + name + ".tris");
    const std::vector<char> want = slurp(dir + "/" + name + ".mask");
    const size_t n = (size_t)nx * ny * nz;
    if (tris.size() != (size_t)ntris * 36 || want.size() != n) {
      std::printf("%s: bad input sizes\n", name.c_str());
      return 1;
    }
    std::vector<uint8_t> got(n);
    int64_t counts[3];
    if (wt_voxelize((const float*)tris.data(), ntris, nx, ny, nz, ground, got.data(), counts) != 0) {
      std::printf("%s: wt_voxelize failed\n", name.c_str());
      return 1;
    }
    if (std::memcmp(got.data(), want.data(), n) != 0 || counts[0] != count || counts[1] != frontal ||
        counts[2] != planform) {
      std::printf("%s: mask or counts differ from the golden case\n", name.c_str());
      return 1;
    }
    std::printf("voxelize %s %dx%dx%d ok\n", name.c_str(), nx, ny, nz);
    ncase++;
  }
  if (ncase == 0) {
    std::printf("no voxelizer cases read\n");
    return 1;
  }
  // analytic bodies: ellipsoids reaching past every face (clipped) and a thin one
  const double ell[] = {10.0, 6.0, 7.0, 14.0, 9.0, 9.0,  30.0, 1.0, 7.0, 5.0, 3.0, 2.0,  20.0, 7.0, 7.0, 0.4, 0.4, 6.0};
  for (int ground = 0; ground < 2; ground++) {
    std::vector<uint8_t> m(40 * 14 * 15);
    int64_t counts[3];
    if (wt_ellipsoids(ell, 3, 40, 14, 15, ground, m.data(), counts) != 0) {
      std::printf("wt_ellipsoids failed\n");
      return 1;
    }
    std::printf("ellipsoids ground=%d cells=%lld ok\n", ground, (long long)counts[0]);
  }
  return 0;
}

template <typename E>
static int oracle_flow(int form) {
  const int nx = 26, ny = 12, nz = 14;
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<uint8_t> solid(n, 0);
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const double dx = x - 8.0, dy = y - 6.0, dz = z - 7.0;
        if (dx * dx + dy * dy + dz * dz <= 9.0) solid[x + nx * (y + ny * z)] = 1;
        if (y == 0) solid[x + nx * (y + ny * z)] = 2;                      // moving belt
      }
  std::vector<E> f0(19 * n), f1(19 * n);
  const size_t real = form == 3 ? 8 : 4;
  std::vector<char> mac(4 * n * real), mac2(4 * n * real);
  oracle_init(form, f0.data(), mac.data(), n, 0.05f);
  f1 = f0;
  double mass = 0.0;
  for (int k = 0; k < 6; k++) {
    oracle_step(form, (k & 1) ? f1.data() : f0.data(), (k & 1) ? f0.data() : f1.data(), solid.data(), mac.data(),
                mac2.data(), nx, ny, nz, 0.05f, 0.52f, 0.17f, &mass);
    mac.swap(mac2);
  }
  double F[3];
  oracle_force(form, f0.data(), solid.data(), nx, ny, nz, F);
  if (!std::isfinite(mass) || !std::isfinite(F[0])) {
    std::printf("form %d: non-finite mass or force\n", form);
    return 1;
  }
  std::printf("oracle form %d: 6 steps, mass %.6f, Fx %.6g ok\n", form, mass, F[0]);
  if (form != 0) return 0;
  // the FP32 field feeds the particle, slice and Cp restatements
  const float* m = (const float*)mac.data();
  std::vector<uint64_t> links(n);
  oracle_link_classes(solid.data(), nx, ny, nz, links.data());
  const int np = 300;
  std::vector<float> pos(4 * np), spd(np, 1.0f);
  for (int i = 0; i < np; i++) {   // spread over and beyond the tunnel, so kill rules fire
    pos[4 * i] = -1.0f + 0.1f * (float)(i % 290);
    pos[4 * i + 1] = 0.2f + 0.04f * (float)(i % 300);
    pos[4 * i + 2] = 0.3f + 0.05f * (float)((7 * i) % 290);
    pos[4 * i + 3] = (float)(i % 50);
  }
  for (int k = 0; k < 5; k++)
    oracle_advect(pos.data(), spd.data(), m, solid.data(), np, nx, ny, nz, 1.0f, 0.05f, 40.0f, 9u + k, 2.0f, 6.0f,
                  7.0f, 4.5f, 5.5f);
  std::vector<uint16_t> q16(4 * np);
  std::vector<uint8_t> q8(np);
  oracle_pack_particles(pos.data(), spd.data(), q16.data(), q8.data(), np, nx, ny, nz);
  for (int axis = 0; axis < 3; axis++)
    for (int field = 0; field < 4; field++) {
      const int dim = axis == 2 ? nz : (axis == 1 ? ny : nx);
      std::vector<uint8_t> out(2 * nx * ny * nz);
      for (int index : {0, dim / 2, dim - 1})
        oracle_slice(m, solid.data(), out.data(), nx, ny, nz, axis, index, field, 0.05f, 16.0f, -1.0f, 2.0f);
    }
  std::vector<float> verts, norms;
  for (int i = 0; i < 40; i++) {   // around the sphere and at the domain's edges
    const double a = 0.157 * i;
    verts.insert(verts.end(), {(float)(8 + 3 * std::cos(a)), (float)(6 + 3 * std::sin(a)), 7.0f});
    norms.insert(norms.end(), {(float)std::cos(a), (float)std::sin(a), 0.0f});
  }
  verts.insert(verts.end(), {0.0f, 0.0f, 0.0f, 25.9f, 11.9f, 13.9f});
  norms.insert(norms.end(), {-1.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f});
  std::vector<int8_t> cp(verts.size() / 3);
  oracle_cp(m, solid.data(), verts.data(), norms.data(), cp.data(), (int64_t)cp.size(), nx, ny, nz, 0.05f, 1.2f);
  std::printf("oracle particles, slices, Cp ok\n");
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("usage: sanitize_main <case dir>\n");
    return 2;
  }
  oracle_set_threads(2);
  if (int e = voxel_cases(argv[1])) return e;
  if (oracle_flow<float>(0) || oracle_flow<uint16_t>(1) || oracle_flow<uint16_t>(2) || oracle_flow<double>(3))
    return 1;
  std::printf("sanitize ok\n");
  return 0;
}
