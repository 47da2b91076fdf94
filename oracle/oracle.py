"""Python face of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, never by the
product package. See oracle/wt_oracle.cpp for what each function restates
(server/lbm.py line references) and for the deliberate choices where the
reference is nondeterministic.

Parity unpinned: the respawn random stream of advect (replaced; checked by
distribution) and Warp's last-ulp FP32 rounding, which cannot be observed
without the reference's runtime (warp-lang 1.17.0, not installed).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

FP32, FP16, BF16, FP64 = 0, 1, 2, 3
ELEM = {FP32: np.float32, FP16: np.uint16, BF16: np.uint16, FP64: np.float64}
REAL = {FP32: np.float32, FP16: np.float32, BF16: np.float32, FP64: np.float64}
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
              (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
              (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)])
OPP = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]

_lib = None
vp, i32, i64, f32, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_uint32


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b  # noqa
            _b.build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_init.argtypes = [i32, vp, vp, i64, f32]
        L.oracle_step.argtypes = [i32, vp, vp, vp, vp, vp, i32, i32, i32, f32, f32, f32, ctypes.POINTER(ctypes.c_double)]
        L.oracle_force.argtypes = [i32, vp, vp, i32, i32, i32, ctypes.POINTER(ctypes.c_double)]
        L.oracle_link_classes.argtypes = [vp, i32, i32, i32, vp]
        L.oracle_advect.argtypes = [vp, vp, vp, vp, i64, i32, i32, i32, f32, f32, f32, u32, f32, f32, f32, f32, f32]
        L.oracle_pack_particles.argtypes = [vp, vp, vp, vp, i64, i32, i32, i32]
        L.oracle_slice.argtypes = [vp, vp, vp, i32, i32, i32, i32, i32, i32, f32, f32, f32, f32]
        L.oracle_cp.argtypes = [vp, vp, vp, vp, vp, i64, i32, i32, i32, f32, f32]
        L.oracle_rng_init.argtypes = [u32, u32]
        L.oracle_rng_init.restype = u32
        L.oracle_rng_f.argtypes = [ctypes.POINTER(u32)]
        L.oracle_rng_f.restype = f32
        L.oracle_f2h.argtypes = [f32]
        L.oracle_f2h.restype = ctypes.c_uint16
        L.oracle_h2f.argtypes = [ctypes.c_uint16]
        L.oracle_h2f.restype = f32
        L.oracle_f2bf.argtypes = [f32]
        L.oracle_f2bf.restype = ctypes.c_uint16
        L.oracle_bf2f.argtypes = [ctypes.c_uint16]
        L.oracle_bf2f.restype = f32
        L.oracle_clamp_hits.argtypes = [ctypes.POINTER(ctypes.c_int64)]
        L.oracle_set_threads.argtypes = [i32]
        L.oracle_get_threads.restype = i32
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(vp)


def clamp_hits():
    """(velocity clamps, face density clamps) counted by every oracle step since the
    last call (lbm.py:130-135, lbm.py:88)."""
    out = (ctypes.c_int64 * 2)()
    lib().oracle_clamp_hits(out)
    return int(out[0]), int(out[1])


def set_threads(n):
    lib().oracle_set_threads(int(n))


class OracleTunnel:
    """CPU state of one tunnel, stepped by the oracle (lbm.py:398-482 semantics)."""

    def __init__(self, dims, form=FP32):
        self.nx, self.ny, self.nz = (int(d) for d in dims)
        self.n = self.nx * self.ny * self.nz
        self.form = form
        self.f = [np.zeros(19 * self.n, ELEM[form]) for _ in range(2)]
        self.mac = np.zeros((self.n, 4), REAL[form])
        self.mac_next = np.zeros((self.n, 4), REAL[form])
        self.solid = np.zeros(self.n, np.uint8)
        self.cur = 0
        self.mass = None

    def init(self, uin):
        lib().oracle_init(self.form, _p(self.f[0]), _p(self.mac), self.n, float(uin))
        self.f[1][:] = self.f[0]
        self.cur = 0

    def set_solid(self, mask):
        self.solid[:] = np.asarray(mask, np.uint8).reshape(-1)

    def step(self, nsteps, uin, tau, smag):
        m = ctypes.c_double(0.0)
        for _ in range(int(nsteps)):
            lib().oracle_step(self.form, _p(self.f[self.cur]), _p(self.f[1 - self.cur]), _p(self.solid),
                              _p(self.mac), _p(self.mac_next), self.nx, self.ny, self.nz,
                              float(uin), float(tau), float(smag), ctypes.byref(m))
            self.mac, self.mac_next = self.mac_next, self.mac
            self.cur ^= 1
        self.mass = m.value

    def device_mac(self):
        """The macroscopic field as the device stores it: float32 (ux, uy, uz, rho)
        for the FP32/FP64 forms; for FP16/BF16 float16 (ux, uy, uz, rho - 1), each
        rounded to nearest even from float32, rho - 1 formed in float32."""
        m = self.mac.astype(np.float32)
        if self.form in (FP16, BF16):
            q = m.copy()
            q[:, 3] = m[:, 3] - np.float32(1.0)
            return q.astype(np.float16)
        return m

    def field(self):
        """(ux, uy, uz, rho) float32 as the device's particle, slice and Cp kernels
        read it (for FP16/BF16: the float16 values widened, rho = 1 + (rho - 1) in
        float32)."""
        return device_field(self.device_mac())

    def forces(self):
        out = (ctypes.c_double * 3)()
        lib().oracle_force(self.form, _p(self.f[self.cur]), _p(self.solid), self.nx, self.ny, self.nz, out)
        return np.array(list(out))

    def raw(self):
        return self.f[self.cur].reshape(19, self.n)

    def populations(self, dtype=np.float64):
        raw = self.raw()
        if self.form in (FP32, FP64):
            return raw.astype(dtype)
        if self.form == FP16:
            g = raw.view(np.float16).astype(np.float32)
        else:
            g = (raw.astype(np.uint32) << 16).view(np.float32)
        return (W.astype(np.float32)[:, None] + g).astype(dtype)


def device_field(mac):
    """float32 (ux, uy, uz, rho) from a device-layout macroscopic field (float32, or
    float16 with rho - 1); advect, slice_ and cp take either layout through it."""
    mac = np.asarray(mac).reshape(-1, 4)
    if mac.dtype == np.float16:
        f = mac.astype(np.float32)
        f[:, 3] = np.float32(1.0) + f[:, 3]
        return f
    return mac.astype(np.float32)


def link_classes(solid, dims):
    nx, ny, nz = dims
    solid = np.ascontiguousarray(solid, np.uint8)
    out = np.zeros(nx * ny * nz, np.uint64)
    lib().oracle_link_classes(_p(solid), nx, ny, nz, _p(out))
    return out


def advect(pos, spd, mac, solid, dims, dt, uin, max_age, seed, box):
    pos = np.array(pos, np.float32, copy=True).reshape(-1, 4)
    spd = np.array(spd, np.float32, copy=True).reshape(-1)
    mac = np.ascontiguousarray(device_field(mac))
    solid = np.ascontiguousarray(solid, np.uint8)
    lib().oracle_advect(_p(pos), _p(spd), _p(mac), _p(solid), pos.shape[0], *dims, float(dt), float(uin),
                        float(max_age), int(seed) & 0xffffffff, *(float(b) for b in box))
    return pos, spd


def pack_particles(pos, spd, dims):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
    spd = np.ascontiguousarray(spd, np.float32)
    q16 = np.zeros(4 * pos.shape[0], np.uint16)
    q8 = np.zeros(pos.shape[0], np.uint8)
    lib().oracle_pack_particles(_p(pos), _p(spd), _p(q16), _p(q8), pos.shape[0], *dims)
    return q16, q8


def slice_dims(dims, axis):
    nx, ny, nz = dims
    return {2: (nx, ny), 1: (nx, nz), 0: (nz, ny)}[axis]


def slice_(mac, solid, dims, axis, index, field, uin, lref, vmin, vmax):
    w, h = slice_dims(dims, axis)
    out = np.zeros(2 * w * h, np.uint8)
    lib().oracle_slice(_p(np.ascontiguousarray(device_field(mac))), _p(np.ascontiguousarray(solid, np.uint8)),
                       _p(out), *dims, axis, index, field, float(uin), float(lref), float(vmin), float(vmax))
    return out


def cp(mac, solid, verts, norms, dims, uin, cp_range):
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    norms = np.ascontiguousarray(norms, np.float32).reshape(-1, 3)
    out = np.zeros(verts.shape[0], np.int8)
    lib().oracle_cp(_p(np.ascontiguousarray(device_field(mac))), _p(np.ascontiguousarray(solid, np.uint8)),
                    _p(verts), _p(norms), _p(out), verts.shape[0], *dims, float(uin), float(cp_range))
    return out


def f2h(x):
    return lib().oracle_f2h(float(x))


def f2bf(x):
    return lib().oracle_f2bf(float(x))
