// Host-side geometry: triangle-soup voxelizer and analytic ellipsoid unions.
//
// wt_voxelize restates web/js/voxelize.js:5-115 in IEEE double precision
// (JavaScript numbers), compiled with -ffp-contract=off so each operation
// rounds exactly as the JavaScript engine's does:
//   1) +x ray parity per (y, z) column, rays jittered by (0.01373, 0.00717)
//      (voxelize.js:9-50);
//   2) a rasterized surface shell, (s+1)(s+2)/2 barycentric samples per
//      triangle (voxelize.js:52-71);
//   3) flood fill from the non-solid boundary cells; cells it cannot reach
//      become solid (voxelize.js:73-97);
//   4) frontal and planform areas (voxelize.js:99-110);
//   5) the y = 0 layer tagged 2 (moving belt) when ground is on (voxelize.js:112).
// The flood fill's result is the set of cells reachable from the boundary,
// which does not depend on the visiting order, so an explicit stack is used.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "wt.h"
#include "wt_error.h"

namespace wt {

// JavaScript Math.round: the nearest integer, ties toward +infinity.
static inline double js_round(double x) {
  double r = std::floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}

// Flood fill of non-marked cells from the domain boundary, then areas and the
// ground tag (voxelize.js:73-114). `vox` holds the marked cells on entry;
// `out` receives the final mask. Unreached open cells become solid.
//
// The fill works on runs: maximal x-intervals of open cells in one (y, z) row.
// Two runs are 6-connected exactly when they lie in rows adjacent in y or z and
// their x-intervals overlap (runs of one row are separated by marked cells), so
// a breadth-first search over runs reaches the same cells as one over cells,
// touching each cell only to build the runs and to write the result.
static void close_and_count(const std::vector<uint8_t>& vox, int nx, int ny, int nz, int ground,
                            uint8_t* out, int64_t counts[3]) {
  const int64_t nrows = (int64_t)ny * nz;
  std::vector<int64_t> row_start(nrows + 1, 0);
  std::vector<int32_t> rx0, rx1;
  rx0.reserve(nrows * 2);
  rx1.reserve(nrows * 2);
  for (int64_t r = 0; r < nrows; r++) {
    row_start[r] = (int64_t)rx0.size();
    const uint8_t* v = vox.data() + r * nx;
    int x = 0;
    while (x < nx) {
      while (x < nx && v[x]) x++;
      if (x >= nx) break;
      const int a = x;
      while (x < nx && !v[x]) x++;
      rx0.push_back(a);
      rx1.push_back(x - 1);
    }
  }
  row_start[nrows] = (int64_t)rx0.size();
  const int64_t nruns = (int64_t)rx0.size();
  std::vector<uint8_t> reached(nruns, 0);
  std::vector<int64_t> queue;
  queue.reserve(1 << 16);
  // seeds: every run in a boundary row, and runs touching x = 0 or x = nx - 1
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++) {
      const int64_t r = y + (int64_t)ny * z;
      const bool edge_row = y == 0 || z == 0 || y == ny - 1 || z == nz - 1;
      for (int64_t q = row_start[r]; q < row_start[r + 1]; q++)
        if (edge_row || rx0[q] == 0 || rx1[q] == nx - 1) { reached[q] = 1; queue.push_back(q); }
    }
  // runs of row nr overlapping [a, b]: the first with x1 >= a, then while x0 <= b
  auto visit_row = [&](int64_t nr, int a, int b) {
    int64_t lo = row_start[nr], hi = row_start[nr + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (rx1[mid] < a) lo = mid + 1; else hi = mid;
    }
    for (int64_t q = lo; q < row_start[nr + 1] && rx0[q] <= b; q++)
      if (!reached[q]) { reached[q] = 1; queue.push_back(q); }
  };
  std::vector<int64_t> run_row(nruns);
  for (int64_t r = 0; r < nrows; r++)
    for (int64_t q = row_start[r]; q < row_start[r + 1]; q++) run_row[q] = r;
  for (size_t h = 0; h < queue.size(); h++) {
    const int64_t q = queue[h];
    const int64_t r = run_row[q];
    const int y = (int)(r % ny), z = (int)(r / ny);
    const int a = rx0[q], b = rx1[q];
    if (y > 0) visit_row(r - 1, a, b);
    if (y < ny - 1) visit_row(r + 1, a, b);
    if (z > 0) visit_row(r - ny, a, b);
    if (z < nz - 1) visit_row(r + ny, a, b);
  }
  // mask, cell count and frontal area (rows with a solid cell), row by row
  int64_t count = 0, frontal = 0;
  for (int64_t r = 0; r < nrows; r++) {
    uint8_t* o = out + r * nx;
    std::memset(o, 1, (size_t)nx);
    int64_t open = 0;
    for (int64_t q = row_start[r]; q < row_start[r + 1]; q++)
      if (reached[q]) {
        std::memset(o + rx0[q], 0, (size_t)(rx1[q] - rx0[q] + 1));
        open += rx1[q] - rx0[q] + 1;
      }
    count += nx - open;
    if (open < nx) frontal++;
  }
  // planform area: (x, z) columns with a solid cell at any y
  int64_t planform = 0;
  std::vector<uint8_t> colf((size_t)nx);
  for (int z = 0; z < nz; z++) {
    std::fill(colf.begin(), colf.end(), 0);
    for (int y = 0; y < ny; y++) {
      const uint8_t* o = out + (int64_t)nx * (y + (int64_t)ny * z);
      for (int x = 0; x < nx; x++) colf[x] |= o[x];
    }
    for (int x = 0; x < nx; x++) planform += colf[x];
  }
  if (ground)
    for (int z = 0; z < nz; z++) std::memset(out + (int64_t)nx * ny * z, 2, (size_t)nx);
  counts[0] = count;
  counts[1] = frontal;
  counts[2] = planform;
}

}  // namespace wt

using namespace wt;

extern "C" int wt_voxelize(const float* tris, int64_t ntris, int nx, int ny, int nz, int ground,
                           uint8_t* out_host, int64_t counts[3]) {
  if (!out_host || !counts || (ntris > 0 && !tris)) return fail(-1, "null argument");
  if (nx < 3 || ny < 3 || nz < 3) return fail(-1, "grid dimensions must be at least 3");
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<uint8_t> vox(n, 0);
  std::vector<std::vector<double>> cols((size_t)ny * nz);
  const double OY = 0.01373, OZ = 0.00717;

  // 1) ray parity along +x, voxelize.js:12-50
  for (int64_t t = 0; t < ntris; t++) {
    const float* o = tris + 9 * t;
    const double ax = o[0], ay = o[1], az = o[2];
    const double bx = o[3], by = o[4], bz = o[5];
    const double cx = o[6], cy = o[7], cz = o[8];
    const double y0d = std::max(1.0, std::ceil(std::min(std::min(ay, by), cy) - OY));
    const double y1d = std::min((double)(ny - 2), std::floor(std::max(std::max(ay, by), cy) - OY));
    const double z0d = std::max(1.0, std::ceil(std::min(std::min(az, bz), cz) - OZ));
    const double z1d = std::min((double)(nz - 2), std::floor(std::max(std::max(az, bz), cz) - OZ));
    if (!(y0d <= y1d) || !(z0d <= z1d)) continue;
    const double e1y = by - ay, e1z = bz - az, e2y = cy - ay, e2z = cz - az;
    const double det = e1y * e2z - e2y * e1z;
    if (det == 0) continue;
    const double id = 1 / det;
    const int y0 = (int)y0d, y1 = (int)y1d, z0 = (int)z0d, z1 = (int)z1d;
    for (int k = z0; k <= z1; k++) {
      const double pz = k + OZ - az;
      for (int j = y0; j <= y1; j++) {
        const double py = j + OY - ay;
        const double w1 = (py * e2z - e2y * pz) * id;
        if (w1 < 0) continue;
        const double w2 = (e1y * pz - py * e1z) * id;
        if (w2 < 0 || w1 + w2 > 1) continue;
        cols[(size_t)j + (size_t)ny * k].push_back(ax + w1 * (bx - ax) + w2 * (cx - ax));
      }
    }
  }
  for (size_t c = 0; c < cols.size(); c++) {
    std::vector<double>& l = cols[c];
    if (l.empty()) continue;
    std::sort(l.begin(), l.end());
    const int64_t j = (int64_t)(c % ny), k = (int64_t)(c / ny);
    const int64_t base = (int64_t)nx * (j + (int64_t)ny * k);
    for (size_t i = 0; i + 1 < l.size(); i += 2) {
      const double a = std::max(1.0, std::ceil(l[i]));
      const double b = std::min((double)(nx - 2), std::floor(l[i + 1]));
      for (double x = a; x <= b; x += 1.0) vox[base + (int64_t)x] = 1;
    }
  }

  // 2) surface shell, voxelize.js:52-71
  for (int64_t t = 0; t < ntris; t++) {
    const float* o = tris + 9 * t;
    const double ax = o[0], ay = o[1], az = o[2];
    const double ux = o[3] - ax, uy = o[4] - ay, uz = o[5] - az;
    const double vx = o[6] - ax, vy = o[7] - ay, vz = o[8] - az;
    const double dx = ux - vx, dy = uy - vy, dz = uz - vz;
    const double L = std::sqrt(std::max(std::max(ux * ux + uy * uy + uz * uz, vx * vx + vy * vy + vz * vz),
                                        dx * dx + dy * dy + dz * dz));
    const double sd = std::max(1.0, std::min(256.0, std::ceil(L / 0.5)));
    const int s = (int)sd;
    for (int a = 0; a <= s; a++) {
      for (int b = 0; a + b <= s; b++) {
        const double fa = (double)a / sd, fb = (double)b / sd;
        const double x = js_round(ax + ux * fa + vx * fb);
        const double y = js_round(ay + uy * fa + vy * fb);
        const double z = js_round(az + uz * fa + vz * fb);
        if (x < 1 || y < 1 || z < 1 || x > nx - 2 || y > ny - 2 || z > nz - 2) continue;
        vox[(int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * (int64_t)z)] = 1;
      }
    }
  }

  // 3-5) closure, areas, ground
  close_and_count(vox, nx, ny, nz, ground, out_host, counts);
  return 0;
}

extern "C" int wt_ellipsoids(const double* ell, int k, int nx, int ny, int nz, int ground, uint8_t* out_host,
                             int64_t counts[3]) {
  if (!out_host || !counts || (k > 0 && !ell)) return fail(-1, "null argument");
  if (nx < 3 || ny < 3 || nz < 3) return fail(-1, "grid dimensions must be at least 3");
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<uint8_t> vox(n, 0);
  for (int e = 0; e < k; e++) {
    const double cx = ell[6 * e], cy = ell[6 * e + 1], cz = ell[6 * e + 2];
    const double ax = ell[6 * e + 3], ay = ell[6 * e + 4], az = ell[6 * e + 5];
    if (!(ax > 0 && ay > 0 && az > 0)) return fail(-1, "ellipsoid %d has a non-positive semi-axis", e);
    const int x0 = std::max(1, (int)std::floor(cx - ax) - 1), x1 = std::min(nx - 2, (int)std::ceil(cx + ax) + 1);
    const int y0 = std::max(1, (int)std::floor(cy - ay) - 1), y1 = std::min(ny - 2, (int)std::ceil(cy + ay) + 1);
    const int z0 = std::max(1, (int)std::floor(cz - az) - 1), z1 = std::min(nz - 2, (int)std::ceil(cz + az) + 1);
    for (int z = z0; z <= z1; z++)
      for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) {
          const double dx = ((double)x - cx) / ax, dy = ((double)y - cy) / ay, dz = ((double)z - cz) / az;
          if (dx * dx + dy * dy + dz * dz <= 1.0) vox[x + (int64_t)nx * (y + (int64_t)ny * z)] = 1;
        }
  }
  close_and_count(vox, nx, ny, nz, ground, out_host, counts);
  return 0;
}
