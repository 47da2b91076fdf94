"""Device state of one wind tunnel and thin wrappers over the libwt kernels.

torch provides the device buffers and the stream; every computation is a
libwt call. Buffer layout (see include/wt.h): two row-blocked SoA
population buffers f[0], f[1] (each x-row holds its 19 planes) of 19*n elements, a class byte per cell, the
macroscopic field mac (float32 ux, uy, uz, rho; binary16 ux, uy, uz, rho - 1 for
FP16/BF16 tunnels), two face-layer snapshots, and the FP64 reduction slots.
"""
import ctypes

import numpy as np
import torch

from . import _lib, autotune
from ._lib import Grid, Params, SeedBox, check, stream_handle, FORMS, FORM_NAMES

ELEM_DTYPE = {0: torch.float32, 1: torch.int16, 2: torch.int16, 3: torch.float64}
# the macroscopic field per cell: float32 (ux, uy, uz, rho) for FP32/FP64 tunnels,
# float16 (ux, uy, uz, rho - 1) for FP16/BF16 tunnels (include/wt.h)
MAC_DTYPE = {0: torch.float32, 1: torch.float16, 2: torch.float16, 3: torch.float32}
# position of direction i's plane inside a row block (include/wt.h)
PLANE_POS = [0, 1, 2, 3, 6, 9, 12, 4, 7, 8, 5, 10, 13, 14, 11, 15, 16, 17, 18]
# population weights w_i (server/lbm.py:18)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def buffer_sizes(dims, form):
    lib = _lib.load()
    s = (ctypes.c_int64 * 6)()
    check(lib.wt_buffer_sizes(int(dims[0]), int(dims[1]), int(dims[2]), int(form), s), 'wt_buffer_sizes')
    return list(s)


def state_bytes(dims, form='fp32'):
    """Device bytes one tunnel occupies (admission footprint)."""
    form = FORMS[form] if isinstance(form, str) else int(form)
    lib = _lib.load()
    return check(lib.wt_state_bytes(int(dims[0]), int(dims[1]), int(dims[2]), form), 'wt_state_bytes')


def cache_policy(dims, form='fp32'):
    """Cache policy (0, 1 or 2) of the single-tunnel step for this size and form,
    as the library chooses it (wt_cache_policy)."""
    form = FORMS[form] if isinstance(form, str) else int(form)
    return check(_lib.load().wt_cache_policy(int(dims[0]), int(dims[1]), int(dims[2]), form), 'wt_cache_policy')


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


F32, I32 = (torch.float32,), (torch.int32,)
SLICE_AXES = {2: ('nx', 'ny'), 1: ('nx', 'nz'), 0: ('nz', 'ny')}   # slice_dims, lbm.py:455-456


class Tunnel:
    """One tunnel resident in HBM."""

    def __init__(self, dims, form='fp32', device=None):
        _lib.require_gpu()
        self.lib = _lib.load()
        self.form = FORMS[form] if isinstance(form, str) else int(form)
        self.nx, self.ny, self.nz = (int(d) for d in dims)
        self.n = self.nx * self.ny * self.nz
        self.device = torch.device(device if device is not None else 'cuda')
        sz = buffer_sizes(dims, self.form)
        dev = self.device
        self.f = [torch.empty(sz[0], dtype=torch.uint8, device=dev) for _ in range(2)]
        self.cls = torch.zeros(sz[1], dtype=torch.uint8, device=dev)
        self.mac = torch.zeros((self.n, 4), dtype=MAC_DTYPE[self.form], device=dev)
        self.shell = [torch.empty(sz[3], dtype=torch.uint8, device=dev) for _ in range(2)]
        self.partials = torch.zeros(sz[4] // 8, dtype=torch.float64, device=dev)
        self.sums = torch.zeros(4, dtype=torch.float64, device=dev)
        self.grid = Grid(self.nx, self.ny, self.nz, self.form, 1, 0,
                         (ctypes.c_void_p * 2)(self.f[0].data_ptr(), self.f[1].data_ptr()),
                         self.cls.data_ptr(), self.mac.data_ptr(),
                         (ctypes.c_void_p * 2)(self.shell[0].data_ptr(), self.shell[1].data_ptr()),
                         self.partials.data_ptr(), self.sums.data_ptr())
        self.gp = ctypes.byref(self.grid)
        self.cur = 0
        # launch shape of this tunnel's size class (carried in its grid descriptor)
        self.shape = autotune.shape_for(self.form_name, self.n)

    def _arg(self, t, name, dtypes, shape=None, min_numel=None):
        """Pointer to tensor t after checking it: the kernels take raw pointers and
        sizes and cannot tell a wrong buffer from a right one, so a tensor on
        another device, of another dtype, not contiguous, of another shape or
        too small raises ValueError here instead of reaching the GPU."""
        if not isinstance(t, torch.Tensor):
            raise ValueError(f'{name} must be a torch tensor, got {type(t).__name__}')
        if t.device.type != self.device.type or (self.device.index is not None and
                                                 t.device.index != self.device.index):
            raise ValueError(f'{name} is on {t.device}, the tunnel on {self.device}')
        if t.dtype not in dtypes:
            raise ValueError(f'{name} must be {"/".join(str(d) for d in dtypes)}, got {t.dtype}')
        if not t.is_contiguous():
            raise ValueError(f'{name} must be contiguous')
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise ValueError(f'{name} must have shape {tuple(shape)}, got {tuple(t.shape)}')
        if min_numel is not None and t.numel() < min_numel:
            raise ValueError(f'{name} holds {t.numel()} elements, needs {min_numel}')
        return _ptr(t)

    def _particles(self, pos, spd, **ints):
        """Pointers to pos (n, 4) float32, spd (n,) float32 and the named (n,) int32
        index arrays (None passes through as a null pointer); returns (n, ptrs)."""
        if not isinstance(pos, torch.Tensor) or pos.dim() != 2:
            raise ValueError('pos must be an (n, 4) float32 tensor')
        n = int(pos.shape[0])
        ptrs = [self._arg(pos, 'pos', F32, (n, 4)), self._arg(spd, 'spd', F32, (n,))]
        for name, t in ints.items():
            ptrs.append(None if t is None else self._arg(t, name, I32, (n,)))
        return n, ptrs

    @property
    def shape(self):
        """Launch shape (cells per thread, tile order) of this tunnel's step kernels,
        passed to the library in its grid descriptor with every call."""
        return self.grid.cpt, self.grid.tile_order

    @shape.setter
    def shape(self, shape):
        cpt, order = (int(v) for v in shape)
        if cpt not in (1, 2, 4):
            raise ValueError(f'cells per thread must be 1, 2 or 4, got {cpt}')
        if order & ~7:
            raise ValueError(f'unknown tile order bits {order:#x}')
        self.grid.cpt, self.grid.tile_order = cpt, order

    @property
    def cache_policy(self):
        """The cache policy the library's single-tunnel step uses for this tunnel."""
        return cache_policy((self.nx, self.ny, self.nz), self.form)

    @property
    def form_name(self):
        return FORM_NAMES[self.form]

    @property
    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in
                   (*self.f, self.cls, self.mac, *self.shell, self.partials, self.sums))

    # ---- state -------------------------------------------------------------
    def init(self, uin, stream=None):
        check(self.lib.wt_init(self.gp, float(uin), stream_handle(stream)), 'wt_init')
        self.cur = 0

    def set_solid(self, mask, stream=None):
        mask = np.ascontiguousarray(np.asarray(mask).reshape(-1), dtype=np.uint8)
        if mask.size != self.n:
            raise ValueError(f'solid mask has {mask.size} cells, grid has {self.n}')
        # staged through pinned memory so the host never waits: the pinned block and the
        # device copy are released to their caching allocators only after the stream
        # has used them
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.stream(s):
            dev = torch.from_numpy(mask).pin_memory().to(self.device, non_blocking=True)
            check(self.lib.wt_set_solid(self.gp, _ptr(dev), stream_handle(s)), 'wt_set_solid')

    def set_mac(self, mac):
        """Point later launches at another (n, 4) macroscopic buffer (float32, or
        float16 for FP16/BF16 tunnels). The step kernel only writes mac (face cells
        read the shell snapshot), so a caller may alternate two buffers and read one
        while a step writes the other."""
        self._arg(mac, 'mac', (MAC_DTYPE[self.form],), (self.n, 4))
        self.mac = mac
        self.grid.mac = mac.data_ptr()

    # ---- hot path ----------------------------------------------------------
    def step(self, nsteps, uin, tau, smag, flags=0, stream=None):
        p = Params(float(uin), float(tau), float(smag))
        check(self.lib.wt_step(self.gp, self.cur, int(nsteps), ctypes.byref(p), int(flags),
                               stream_handle(stream)), 'wt_step')
        self.cur ^= int(nsteps) & 1

    def step_advect(self, uin, tau, smag, flags, mac_prev, pos, spd, ids, dt, max_age, seed, box, stream=None):
        """One step fused with the advection (slot storage) of the particles of the
        previous step from mac_prev, the field that step wrote (not self.mac)."""
        p = Params(float(uin), float(tau), float(smag))
        sb = SeedBox(*(float(v) for v in box))
        mp = self._arg(mac_prev, 'mac_prev', (MAC_DTYPE[self.form],), (self.n, 4))
        n, (pp, sp, ip) = self._particles(pos, spd, ids=ids)
        check(self.lib.wt_step_advect(self.gp, self.cur, ctypes.byref(p), int(flags), mp, pp, sp, ip, n, float(dt),
                                      float(uin), float(max_age), int(seed) & 0xffffffff, ctypes.byref(sb),
                                      stream_handle(stream)), 'wt_step_advect')
        self.cur ^= 1

    def forces(self, stream=None):
        """Momentum-exchange force on f[cur] into self.sums[0:3] (device)."""
        check(self.lib.wt_forces(self.gp, self.cur, stream_handle(stream)), 'wt_forces')
        return self.sums

    def link_classes(self, stream=None):
        out = torch.empty(self.n, dtype=torch.int64, device=self.device)
        check(self.lib.wt_link_classes(self.gp, _ptr(out), stream_handle(stream)), 'wt_link_classes')
        return out

    # ---- views ---------------------------------------------------------------
    def macroscopic(self, mac=None):
        """(n, 4) float32 (ux, uy, uz, rho) on the device as the particle, slice and
        Cp kernels read it: FP16/BF16 tunnels store (ux, uy, uz, rho - 1) in float16,
        and rho is rebuilt as 1 + (rho - 1) in float32."""
        m = self.mac if mac is None else mac
        if m.dtype == torch.float16:
            f = m.float()
            f[:, 3] += 1.0
            return f
        return m
    def raw_populations(self, which=None):
        """f[which] (default f[cur]) as its storage dtype, shape (19, n) in the
        reference's plane order (population i of cell idx at [i, idx]), on the
        device. The buffer itself is row-blocked SoA (each x-row holds its 19
        planes in turn); this returns a reordered copy."""
        k = self.cur if which is None else which
        rows = self.f[k].view(ELEM_DTYPE[self.form]).view(self.ny * self.nz, 19, self.nx)
        return rows[:, PLANE_POS, :].permute(1, 0, 2).reshape(19, self.n)

    def load_raw_populations(self, raw, which=None):
        """Write f[which] (default both buffers) from a (19, n) array of storage elements
        in the reference's plane order, the inverse of raw_populations(). Solid cells
        keep whatever this writes, as a step never stores them."""
        dt = ELEM_DTYPE[self.form]
        src = torch.as_tensor(np.ascontiguousarray(raw).reshape(19, self.n).view(
            {torch.float32: np.float32, torch.int16: np.int16, torch.float64: np.float64}[dt]))
        src = src.to(self.device).view(19, self.ny * self.nz, self.nx).permute(1, 0, 2)
        for k in ((0, 1) if which is None else (which,)):
            rows = self.f[k].view(dt).view(self.ny * self.nz, 19, self.nx)
            rows[:, PLANE_POS, :] = src

    def populations(self, which=None, dtype=np.float64):
        """f_i of every cell on the host, shape (19, n), reconstructed from the storage form."""
        raw = self.raw_populations(which).cpu().numpy()
        if self.form == 0:
            return raw.astype(dtype)
        if self.form == 3:
            return raw.astype(dtype)
        u16 = raw.view(np.uint16)
        if self.form == 1:
            g = u16.view(np.float16).astype(np.float32)
        else:
            g = (u16.astype(np.uint32) << 16).view(np.float32)
        return (W.astype(np.float32)[:, None] + g).astype(dtype)

    # ---- checkpoint / resume ---------------------------------------------------
    def checkpoint(self, stream=None):
        """Copy the state that defines the next step -- f[cur], the class bytes, mac
        and the face-layer snapshot of the current parity -- into pinned host memory
        (asynchronous on `stream`; synchronize before reading or freeing the tunnel)."""
        k = self.cur
        st = {'dims': (self.nx, self.ny, self.nz), 'form': self.form}
        for name, t in (('f', self.f[k]), ('cls', self.cls), ('mac', self.mac), ('shell', self.shell[k])):
            h = torch.empty(t.shape, dtype=t.dtype).pin_memory()
            h.copy_(t, non_blocking=True)
            st[name] = h
        return st

    @classmethod
    def from_checkpoint(cls_, st, device=None, stream=None):
        """A new resident tunnel continuing exactly where checkpoint() left off."""
        t = cls_(st['dims'], st['form'], device)
        t.f[0].copy_(st['f'], non_blocking=True)
        # solid cells are never written by a step, so both buffers hold the same
        # values there; restoring both keeps the raw state identical
        t.f[1].copy_(st['f'], non_blocking=True)
        t.cls.copy_(st['cls'], non_blocking=True)
        t.mac.copy_(st['mac'], non_blocking=True)
        t.shell[0].copy_(st['shell'], non_blocking=True)
        t.cur = 0
        return t

    # ---- particles, slices, surface -------------------------------------------
    def advect(self, pos, spd, dt, uin, max_age, seed, box, stream=None):
        sb = SeedBox(*(float(v) for v in box))
        n, (pp, sp) = self._particles(pos, spd)
        check(self.lib.wt_advect(self.gp, pp, sp, n, float(dt), float(uin), float(max_age), int(seed) & 0xffffffff,
                                 ctypes.byref(sb), stream_handle(stream)), 'wt_advect')

    def advect_ordered(self, pos, spd, order, dt, uin, max_age, seed, box, stream=None):
        sb = SeedBox(*(float(v) for v in box))
        n, (pp, sp, op) = self._particles(pos, spd, order=order)
        check(self.lib.wt_advect_ordered(self.gp, pp, sp, op, n, float(dt), float(uin), float(max_age),
                                         int(seed) & 0xffffffff, ctypes.byref(sb), stream_handle(stream)),
              'wt_advect_ordered')

    def advect_slots(self, pos, spd, ids, dt, uin, max_age, seed, box, stream=None):
        """Advection of slot storage: slot j holds particle ids[j] (int32 tensor)."""
        sb = SeedBox(*(float(v) for v in box))
        n, (pp, sp, ip) = self._particles(pos, spd, ids=ids)
        check(self.lib.wt_advect_slots(self.gp, pp, sp, ip, n, float(dt), float(uin), float(max_age),
                                       int(seed) & 0xffffffff, ctypes.byref(sb), stream_handle(stream)),
              'wt_advect_slots')

    def particle_gather(self, pos, spd, ids, order, pos_out, spd_out, ids_out, stream=None):
        """pos_out[j] = pos[order[j]] (and spd, ids; ids None = identity)."""
        n, (pp, sp, ip, op) = self._particles(pos, spd, ids=ids, order=order)
        _, (po, so, io) = self._particles(pos_out, spd_out, ids_out=ids_out)
        if int(pos_out.shape[0]) != n or io is None:
            raise ValueError('particle_gather outputs must match the inputs (ids_out included)')
        check(self.lib.wt_particle_gather(pp, sp, ip, op, n, po, so, io, stream_handle(stream)),
              'wt_particle_gather')

    def pack_particles_slots(self, pos, spd, ids, q16, q8, stream=None):
        n, (pp, sp, ip) = self._particles(pos, spd, ids=ids)
        check(self.lib.wt_pack_particles_slots(self.gp, pp, sp, ip,
                                               self._arg(q16, 'q16', (torch.int16, torch.uint16), min_numel=4 * n),
                                               self._arg(q8, 'q8', (torch.uint8,), min_numel=n), n,
                                               stream_handle(stream)), 'wt_pack_particles_slots')

    def particle_order(self, pos, order, workspace=None, stream=None):
        """order[:] = particle ids sorted by cell index; returns the workspace for reuse."""
        if not isinstance(pos, torch.Tensor) or pos.dim() != 2:
            raise ValueError('pos must be an (n, 4) float32 tensor')
        n = int(pos.shape[0])
        pp, op = self._arg(pos, 'pos', F32, (n, 4)), self._arg(order, 'order', I32, (n,))
        need = check(self.lib.wt_particle_order_workspace_bytes(n), 'wt_particle_order_workspace_bytes')
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        wp = self._arg(workspace, 'workspace', (torch.uint8,))
        check(self.lib.wt_particle_order(self.gp, pp, n, op, wp, workspace.numel(), stream_handle(stream)),
              'wt_particle_order')
        return workspace

    def pack_particles(self, pos, spd, q16, q8, stream=None):
        n, (pp, sp) = self._particles(pos, spd)
        check(self.lib.wt_pack_particles(self.gp, pp, sp,
                                         self._arg(q16, 'q16', (torch.int16, torch.uint16), min_numel=4 * n),
                                         self._arg(q8, 'q8', (torch.uint8,), min_numel=n), n,
                                         stream_handle(stream)), 'wt_pack_particles')

    def slice_dims(self, axis):
        """(w, h) of a slice perpendicular to axis (lbm.py:455-456)."""
        if int(axis) not in SLICE_AXES:
            raise ValueError(f'slice axis must be 0, 1 or 2, got {axis}')
        a, b = SLICE_AXES[int(axis)]
        return getattr(self, a), getattr(self, b)

    def slice(self, out, axis, index, field, uin, lref, vmin, vmax, stream=None):
        """out[0:2wh] = field bytes and fluid bytes of the plane (slice_kernel, lbm.py:278-318)."""
        w, h = self.slice_dims(axis)
        op = self._arg(out, 'slice output', (torch.uint8,), min_numel=2 * w * h)
        check(self.lib.wt_slice(self.gp, op, int(axis), int(index), int(field), float(uin), float(lref),
                                float(vmin), float(vmax), stream_handle(stream)), 'wt_slice')

    def cp(self, verts, norms, out, uin, cp_range, stream=None):
        if not isinstance(verts, torch.Tensor) or verts.dim() != 2:
            raise ValueError('verts must be an (nv, 3) float32 tensor')
        nv = int(verts.shape[0])
        check(self.lib.wt_cp(self.gp, self._arg(verts, 'verts', F32, (nv, 3)),
                             self._arg(norms, 'norms', F32, (nv, 3)),
                             self._arg(out, 'cp output', (torch.int8,), min_numel=nv), nv, float(uin),
                             float(cp_range), stream_handle(stream)), 'wt_cp')
