// Writes tests/golden/voxel_golden.json + voxel_golden.bin: triangle soups and the
// masks/counts the reference voxelizer (web/js/voxelize.js) returns for them.
//
// Run once, in the build container only (Node v12):
//   node tests/golden/make_voxel_golden.js /root/reference/web/js/voxelize.js
// The reference file is read and evaluated in memory; nothing of it is written.
// The fixture holds data only: float32 triangles (lattice coordinates) and the
// zlib-compressed uint8 masks, with counts {count, frontal, planform}.
'use strict';
const fs = require('fs');
const path = require('path');
const zlib = require('zlib');

const src = fs.readFileSync(process.argv[2] || '/root/reference/web/js/voxelize.js', 'utf8');
const voxelize = new Function(src.replace(/export\s+function/g, 'function') + '\nreturn voxelize;')();

// deterministic generator (mulberry32)
function rng(seed) {
  let a = seed >>> 0;
  return () => {
    a = (a + 0x6D2B79F5) >>> 0;
    let t = a;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

function sphere(cx, cy, cz, r, nu, nv, flip) {
  const out = [];
  const P = (i, j) => {
    const th = Math.PI * i / nv, ph = 2 * Math.PI * j / nu;
    return [cx + r * Math.sin(th) * Math.cos(ph), cy + r * Math.cos(th), cz + r * Math.sin(th) * Math.sin(ph)];
  };
  for (let i = 0; i < nv; i++) for (let j = 0; j < nu; j++) {
    const a = P(i, j), b = P(i + 1, j), c = P(i + 1, j + 1), d = P(i, j + 1);
    const t1 = flip ? [a, c, b] : [a, b, c];
    const t2 = flip ? [a, d, c] : [a, c, d];
    out.push(...t1[0], ...t1[1], ...t1[2], ...t2[0], ...t2[1], ...t2[2]);
  }
  return out;
}

function box(x0, y0, z0, x1, y1, z1, rot) {
  const v = [[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0], [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]];
  let pts = v;
  if (rot) {
    const cx = (x0 + x1) / 2, cy = (y0 + y1) / 2, cz = (z0 + z1) / 2;
    const [a, b] = rot;
    pts = v.map(([x, y, z]) => {
      let X = x - cx, Y = y - cy, Z = z - cz;
      const X1 = X * Math.cos(a) - Y * Math.sin(a), Y1 = X * Math.sin(a) + Y * Math.cos(a);
      const X2 = X1 * Math.cos(b) + Z * Math.sin(b), Z2 = -X1 * Math.sin(b) + Z * Math.cos(b);
      return [X2 + cx, Y1 + cy, Z2 + cz];
    });
  }
  const F = [[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]];
  const out = [];
  for (const f of F) for (const k of f) out.push(...pts[k]);
  return out;
}

const R = rng(20251203);
const soup = [];
for (let t = 0; t < 60; t++) {
  const c = [6 + R() * 30, 4 + R() * 18, 4 + R() * 20];
  for (let k = 0; k < 3; k++) soup.push(c[0] + (R() - 0.5) * 9, c[1] + (R() - 0.5) * 9, c[2] + (R() - 0.5) * 9);
}

const cases = [
  { name: 'sphere', dims: [48, 28, 30], ground: false, tris: sphere(20.4, 12.1, 13.7, 7.3, 40, 20, false) },
  { name: 'sphere_ground', dims: [48, 28, 30], ground: true, tris: sphere(18.0, 9.25, 15.0, 8.1, 24, 12, false) },
  { name: 'hollow_shell', dims: [40, 32, 32], ground: false,
    tris: sphere(19.3, 15.6, 15.9, 11.2, 36, 18, false).concat(sphere(19.3, 15.6, 15.9, 6.4, 28, 14, true)) },
  { name: 'box', dims: [36, 20, 24], ground: true, tris: box(7.3, 1.0, 6.6, 21.8, 9.4, 17.05, null) },
  { name: 'tilted_box', dims: [40, 30, 30], ground: false, tris: box(12.2, 8.3, 9.1, 27.7, 19.9, 20.6, [0.37, -0.61]) },
  { name: 'thin_plate', dims: [32, 24, 24], ground: false,
    tris: [8.5, 3.2, 4.1, 8.5, 20.7, 4.1, 8.5, 20.7, 19.6, 8.5, 3.2, 4.1, 8.5, 20.7, 19.6, 8.5, 3.2, 19.6] },
  { name: 'big_triangle', dims: [200, 24, 24], ground: false,
    tris: [1.5, 2.5, 2.5, 198.2, 3.1, 20.4, 3.3, 21.7, 11.9] },
  { name: 'random_soup', dims: [44, 28, 30], ground: true, tris: soup },
];

const index = [];
const chunks = [];
let off = 0;
for (const c of cases) {
  const tris = new Float32Array(c.tris);
  const r = voxelize(tris, c.dims[0], c.dims[1], c.dims[2], c.ground);
  const tb = Buffer.from(tris.buffer);
  const mb = zlib.deflateSync(Buffer.from(r.solid.buffer));
  index.push({ name: c.name, dims: c.dims, ground: c.ground, ntris: tris.length / 9,
    tris_offset: off, tris_bytes: tb.length, mask_offset: off + tb.length, mask_zbytes: mb.length,
    count: r.count, frontal: r.frontal, planform: r.planform });
  chunks.push(tb, mb);
  off += tb.length + mb.length;
}
const dir = __dirname;
fs.writeFileSync(path.join(dir, 'voxel_golden.bin'), Buffer.concat(chunks));
fs.writeFileSync(path.join(dir, 'voxel_golden.json'), JSON.stringify({
  generator: 'tests/golden/make_voxel_golden.js', reference: 'web/js/voxelize.js', node: process.version,
  cases: index }, null, 1) + '\n');
console.log(index.map(c => `${c.name}: count=${c.count} frontal=${c.frontal} planform=${c.planform}`).join('\n'));
