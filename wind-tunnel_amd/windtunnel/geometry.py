"""Solid masks: the voxelize.js restatement and analytic bodies (host code in libwt).

Masks are uint8 arrays of n = nx*ny*nz cells, x fastest (server/lbm.py:78-80):
0 fluid, 1 static solid, 2 moving belt (the y = 0 layer when ground is on,
web/js/voxelize.js:112).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


def voxelize(tris, dims, ground):
    """Triangle soup in lattice coordinates (T x 9 float32) -> (mask, counts).

    counts = {'count', 'frontal', 'planform'} as voxelize.js:114 returns them.
    """
    tris = np.ascontiguousarray(np.asarray(tris, dtype=np.float32).reshape(-1, 9))
    nx, ny, nz = (int(d) for d in dims)
    out = np.zeros(nx * ny * nz, dtype=np.uint8)
    c = (ctypes.c_int64 * 3)()
    check(_lib.load().wt_voxelize(tris.ctypes.data_as(ctypes.c_void_p), tris.shape[0], nx, ny, nz,
                                  int(bool(ground)), out.ctypes.data_as(ctypes.c_void_p), c), 'wt_voxelize')
    return out, {'count': c[0], 'frontal': c[1], 'planform': c[2]}


def ellipsoids(ells, dims, ground=False):
    """Union of ellipsoids (rows cx, cy, cz, ax, ay, az), cell centres inside if
    sum(((p - c) / a)^2) <= 1, clipped to [1, dim - 2], closed by the voxelize.js
    flood fill, belt layer when ground is on. Returns (mask, counts)."""
    ells = np.ascontiguousarray(np.asarray(ells, dtype=np.float64).reshape(-1, 6))
    nx, ny, nz = (int(d) for d in dims)
    out = np.zeros(nx * ny * nz, dtype=np.uint8)
    c = (ctypes.c_int64 * 3)()
    check(_lib.load().wt_ellipsoids(ells.ctypes.data_as(ctypes.c_void_p), ells.shape[0], nx, ny, nz,
                                    int(bool(ground)), out.ctypes.data_as(ctypes.c_void_p), c), 'wt_ellipsoids')
    return out, {'count': c[0], 'frontal': c[1], 'planform': c[2]}


def sphere(dims, centre, diameter, ground=False):
    """Sphere of the given diameter voxelized at cell centres (<= r^2 rule)."""
    r = diameter / 2.0
    return ellipsoids([[centre[0], centre[1], centre[2], r, r, r]], dims, ground)


def car_body(dims, seed, count=12, box=None, axes=(6.0, 40.0)):
    """Seeded union of `count` ellipsoids resting on the ground plane: semi-axes
    uniform in `axes`, centres uniform in `box` = ((x0, x1), (y0, y1), (z0, z1))
    (default a 200 x 80 x 60 box starting at x = 100, y = 1, centred in z)."""
    nx, ny, nz = dims
    if box is None:
        zc = nz / 2.0
        box = ((100.0, 300.0), (1.0, 81.0), (zc - 30.0, zc + 30.0))
    rng = np.random.default_rng(seed)
    ells = []
    for _ in range(count):
        a = rng.uniform(axes[0], axes[1], size=3)
        c = [rng.uniform(box[0][0], box[0][1]), rng.uniform(box[1][0], box[1][1]),
             rng.uniform(box[2][0], box[2][1])]
        ells.append([c[0], c[1], c[2], a[0], a[1], a[2]])
    return np.array(ells)
