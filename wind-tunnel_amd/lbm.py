"""Drop-in replacement for the reference solver module server/lbm.py.

`from lbm import BYTES_PER_CELL, RESOLUTIONS, Solver` (server/app.py:15) keeps
working: the names, constructor, methods, attributes and the frame() dict are
those of server/lbm.py:21-30, 395-482. Underneath, every kernel is a HIP
kernel of libwt.so (gfx950) called through its C ABI; torch only allocates
the device buffers. There is no Warp and no CPU path: without a HIP device
the constructor raises.

Differences from the reference, all deliberate:
  * Face cells read the previous step's macroscopic field (deterministic;
    the reference races, lbm.py:87 vs 166).
  * Forces are reduced in FP64 in a fixed order (lbm.py:195-197 uses FP32
    atomics), and are fused into the last step of Solver.step.
  * The particle respawn stream is a documented counter-based hash instead of
    Warp's rand_init/randf (lbm.py:256-259).
  * Optional keyword `storage` ('fp32' default, 'fp16', 'bf16', 'fp64') and
    `seed` (for set_particles; the reference's generator is unseeded,
    lbm.py:430).
  * Particles live in cell-ordered slots (re-sorted every `sort_every`
    advections), each slot carrying its particle id; frame() still returns
    them in id order and every particle's values are unchanged.
  * step(n) stores the macroscopic field only in its last lattice step, the
    one whose field the advection and frame() read (lbm.py:443-453); face
    cells read their own snapshot, so no other result depends on it.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__))) if __package__ in (None, '') else None

import torch  # noqa: E402

from windtunnel import _lib  # noqa: E402
from windtunnel.tunnel import Tunnel, state_bytes  # noqa: E402

# Fixed grids, lbm.py:21-29
RESOLUTIONS = {
    'test': (64, 28, 36),
    'low': (128, 56, 72),
    'medium': (176, 80, 96),
    'high': (256, 112, 136),
    'ultra': (352, 144, 176),
    'extreme': (448, 192, 224),
    'max': (544, 224, 272),
}
# Reference footprint per cell with FP32 populations, lbm.py:30. Exact sizes
# per storage form come from footprint().
BYTES_PER_CELL = 19 * 4 * 2 + 16 + 1
CP_RANGE = 1.2  # lbm.py:395


def footprint(dims, storage='fp32'):
    """Exact device bytes of one Solver's tunnel state for admission."""
    return state_bytes(dims, storage)


class Device:
    """The few device facts app.py reads from Warp (app.py:56-59)."""

    def __init__(self, index=0):
        self.is_cuda = torch.cuda.is_available()
        self.index = index
        if self.is_cuda:
            prop = torch.cuda.get_device_properties(index)
            self.name = prop.name
            self.total_memory = int(prop.total_memory)
        else:
            self.name = 'CPU (no HIP device found)'
            self.total_memory = 2 * 1024 ** 3

    def __str__(self):
        return f'cuda:{self.index}' if self.is_cuda else 'cpu'


def get_preferred_device():
    return Device(0)


def synchronize():
    if torch.cuda.is_available():
        torch.cuda.synchronize()


class Solver:
    """Same contract as server/lbm.py:398-482."""

    def __init__(self, dims, device=None, storage='fp32', seed=None):
        self.dev = device
        self.nx, self.ny, self.nz = (int(d) for d in dims)
        self.n = self.nx * self.ny * self.nz
        self.T = Tunnel(dims, storage)
        self.storage = storage
        self.solid = self.T.cls          # device class bytes (low 2 bits: solid value)
        self.mac = self.T.mac            # device (ux, uy, uz, rho): float4, or float16 with rho - 1 (fp16/bf16)
        self.slice_max = max(self.nx * self.ny, self.nx * self.nz, self.nz * self.ny)
        self.slice_buf = torch.zeros(self.slice_max * 2, dtype=torch.uint8, device='cuda')
        self.num_p = 0
        self.num_verts = 0
        self.seed = 1
        self._rng_seed = seed
        self._forces_fresh = False
        self.sort_every = 8       # advections between re-sorts of the particle slots
        self._advects = 0
        self._order_ws = None

    @property
    def cur(self):
        return self.T.cur

    # lbm.py:415-417
    def reset(self, uin):
        self.T.init(uin)
        self._forces_fresh = False

    # lbm.py:419-421
    def set_solid(self, arr):
        assert arr.size == self.n
        self.T.set_solid(np.asarray(arr).astype(np.uint8).reshape(-1))
        self._forces_fresh = False

    # lbm.py:423-427
    def set_surface(self, verts, norms):
        verts = np.ascontiguousarray(np.asarray(verts, dtype=np.float32).reshape(-1, 3))
        norms = np.ascontiguousarray(np.asarray(norms, dtype=np.float32).reshape(-1, 3))
        self.num_verts = len(verts)
        self.verts = torch.from_numpy(verts).cuda()
        self.norms = torch.from_numpy(norms).cuda()
        self.cp_buf = torch.zeros(max(1, self.num_verts), dtype=torch.int8, device='cuda')

    # lbm.py:429-440
    def set_particles(self, count, p):
        rng = np.random.default_rng(self._rng_seed)
        d = np.zeros((count, 4), np.float32)
        d[:, 0] = 2 + rng.random(count) * (self.nx - 8)
        d[:, 1] = p['seedCy'] + (rng.random(count) * 2 - 1) * p['seedHy']
        d[:, 2] = p['seedCz'] + (rng.random(count) * 2 - 1) * p['seedHz']
        d[:, 3] = rng.random(count) * p['maxAge'] * 0.5
        self.num_p = count
        self.pos = torch.from_numpy(d).cuda()
        self.spd = torch.ones(count, dtype=torch.float32, device='cuda')
        self.ids = torch.arange(count, dtype=torch.int32, device='cuda')
        self._spare = (torch.empty_like(self.pos), torch.empty_like(self.spd), torch.empty_like(self.ids))
        self._order = torch.empty(count, dtype=torch.int32, device='cuda')
        self._order_ws = None
        self._advects = 0
        self.q16 = torch.zeros(count * 4, dtype=torch.int16, device='cuda')
        self.q8 = torch.zeros(count, dtype=torch.uint8, device='cuda')

    # lbm.py:442-453
    def step(self, steps, p):
        steps = int(steps)
        if steps > 0:
            # only the last step's macroscopic field is read (by the advection with
            # dt = steps and by frame()), so the steps before it do not store it
            if steps > 1:
                self.T.step(steps - 1, p['uIn'], p['tau'], p['smag'], flags=_lib.WT_FLAG_NO_MAC)
            self.T.step(1, p['uIn'], p['tau'], p['smag'], flags=_lib.WT_FLAG_FORCES)
            self._forces_fresh = True
        self.advance_particles(steps, p)

    def advance_particles(self, steps, p):
        """The particle half of step(): one advection with dt = steps (lbm.py:448-453)."""
        steps = int(steps)
        if self.num_p and steps:
            if self._advects % self.sort_every == 0:
                self._order_ws = self.T.particle_order(self.pos, self._order, self._order_ws)
                pos2, spd2, ids2 = self._spare
                self.T.particle_gather(self.pos, self.spd, self.ids, self._order, pos2, spd2, ids2)
                self._spare = (self.pos, self.spd, self.ids)
                self.pos, self.spd, self.ids = pos2, spd2, ids2
            self._advects += 1
            self.seed += 1
            self.T.advect_slots(self.pos, self.spd, self.ids, float(steps), p['uIn'], p['maxAge'], self.seed,
                                (p['seedX'], p['seedCy'], p['seedCz'], p['seedHy'], p['seedHz']))

    def set_forces(self, sums):
        """Take the force of the current state from a batched step's FP64 sums
        (device tensor of 4: fx, fy, fz, mass) instead of recomputing it."""
        self.T.sums.copy_(sums)
        self._forces_fresh = True

    # lbm.py:455-456
    def slice_dims(self, axis):
        return {2: (self.nx, self.ny), 1: (self.nx, self.nz), 0: (self.nz, self.ny)}[axis]

    # lbm.py:458-482
    def frame(self, p, want_cp):
        out = {}
        if self.num_p:
            self.T.pack_particles_slots(self.pos, self.spd, self.ids, self.q16, self.q8)
            out['parts16'] = self.q16.cpu().numpy().view(np.uint16)
            out['parts8'] = self.q8.cpu().numpy()
        sl = p.get('slice')
        if sl and sl.get('field', -1) >= 0:
            axis = int(sl['axis'])
            w, h = self.slice_dims(axis)
            dim = (self.nx, self.ny, self.nz)[axis]
            index = max(0, min(dim - 1, int(sl['index'])))
            self.T.slice(self.slice_buf, axis, index, int(sl['field']), p['uIn'], p['lref'],
                         float(sl['vmin']), float(sl['vmax']))
            out['slice'] = self.slice_buf[:w * h * 2].cpu().numpy()
            out['sliceW'], out['sliceH'], out['sliceAxis'] = w, h, axis
        if want_cp and self.num_verts:
            self.T.cp(self.verts, self.norms, self.cp_buf, p['uIn'], CP_RANGE)
            out['cp'] = self.cp_buf.cpu().numpy()
        if not self._forces_fresh:
            self.T.forces()
            self._forces_fresh = True
        out['forces'] = self.T.sums[:3].cpu().numpy().tolist()
        return out
