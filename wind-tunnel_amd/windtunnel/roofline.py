"""Kernel-level measurements on the GPU: achievable HBM bandwidth of a plain
copy, the fused step kernel's time per launch for each storage form, and a
sweep of its launch shapes.

    python -m windtunnel.roofline [--forms fp32,fp16] [--steps 100] [--sweep] [--copy-gb 4]
                                  [--particles N --fused fp32]

Each form runs the launch shape tuned for its size class (windtunnel/tuning.json,
as sessions and bench.py do), or with --sweep a list of shapes. Prints one JSON
object per measurement. Algorithmic bytes per cell of the step:
19 populations read + 19 written, 1 class byte, 16 B macroscopic store
(FP32 169, FP64 321: 16 B float4 macroscopic field; FP16/BF16 85: 2-byte
deviations and an 8 B binary16 macroscopic field). With --particles, the forms named by
--fused are timed as wt_step_advect launches (the step fused with the slot
advection of N particles seeded on the inlet plane), priced at 40 more bytes
per particle: id 4, position 16 read + 16 written, speed 4 written.
"""
import argparse
import ctypes
import json
import sys

import numpy as np
import torch

from . import _lib, autotune, geometry
from .tunnel import Tunnel

BYTES_PER_CELL = {'fp32': 169, 'fp16': 85, 'bf16': 85, 'fp64': 321}
BYTES_PER_PARTICLE = 40
PEAK_GBS = 8000.0


def copy_bandwidth(gbytes=4.0, reps=20, blocks=0):
    lib = _lib.load()
    nbytes = int(gbytes * 1e9) // 16 * 16
    a = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
    b = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
    a.fill_(1)
    s = _lib.stream_handle()
    for _ in range(3):
        _lib.check(lib.wt_stream_copy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()), nbytes,
                                      blocks, s), 'wt_stream_copy')
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        lib.wt_stream_copy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()), nbytes, blocks, s)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return {'what': 'stream_copy', 'bytes_moved': 2 * nbytes, 'ms': ms, 'gbs': 2 * nbytes / (ms * 1e-3) / 1e9}


def c2_mask(dims=(512, 256, 256), seed=2025):
    ells = geometry.car_body(dims, seed)
    return geometry.ellipsoids(ells, dims, ground=True)[0]


class _Rake:
    """num_p particles on the inlet plane in slot storage sorted by cell, and
    two mac buffers for wt_step_advect (the step writes one while the other is
    advected)."""
    def __init__(self, t, num_p, seed=0):
        from .session import inlet_rake
        dims = (t.nx, t.ny, t.nz)
        self.box = inlet_rake(dims)
        x, cy, cz, hy, hz = self.box
        rng = np.random.default_rng(seed)
        d = np.zeros((num_p, 4), np.float32)
        d[:, 0] = x + rng.random(num_p) * 2.0
        d[:, 1] = cy + (rng.random(num_p) * 2 - 1) * hy
        d[:, 2] = cz + (rng.random(num_p) * 2 - 1) * hz
        pos = torch.from_numpy(d).to(t.device)
        spd = torch.ones(num_p, dtype=torch.float32, device=t.device)
        ids = torch.arange(num_p, dtype=torch.int32, device=t.device)
        # slots in cell order, as a session keeps them (re-sorted every 100 steps)
        order = torch.empty(num_p, dtype=torch.int32, device=t.device)
        t.particle_order(pos, order, None)
        self.pos, self.spd, self.ids = torch.empty_like(pos), torch.empty_like(spd), torch.empty_like(ids)
        t.particle_gather(pos, spd, ids, order, self.pos, self.spd, self.ids)
        self.macs = (t.mac, t.mac.clone())
        self.max_age = 2.2 * t.nx / 0.05
        self.k = 0

    def step(self, t, flags):
        b = self.k & 1
        t.set_mac(self.macs[b])
        t.step_advect(0.05, 0.51, 0.16, flags, self.macs[b ^ 1], self.pos, self.spd, self.ids, 1.0,
                      self.max_age, self.k + 1, self.box)
        self.k += 1


def step_timing(form, dims, mask, steps=100, warmup=10, flags=0, tunnel=None, particles=0, shape=None):
    t = tunnel or Tunnel(dims, form)
    if shape is not None:
        t.shape = shape     # else the tunnel's own (the tuned shape of its size class)
    if tunnel is None:
        t.init(0.05)
        t.set_solid(mask)
    rake = _Rake(t, particles) if particles else None

    def one():
        if rake:
            rake.step(t, flags)
        else:
            t.step(1, 0.05, 0.51, 0.16, flags)
    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for k in range(steps):
        evs[k][0].record()
        one()
        evs[k][1].record()
    torch.cuda.synchronize()
    if rake:
        t.set_mac(rake.macs[0])
    ms = np.array([a.elapsed_time(b) for a, b in evs])
    n = t.n
    med = float(np.median(ms))
    nbytes = BYTES_PER_CELL[form] * n + BYTES_PER_PARTICLE * particles
    return t, {'what': 'step_advect' if particles else 'step', 'form': form, 'dims': list(dims),
               'particles': particles, 'cpt_remap': list(t.shape),
               'flags': flags, 'ms_mean': float(ms.mean()), 'ms_median': med, 'ms_min': float(ms.min()),
               'mlups': n / (med * 1e-3) / 1e6, 'bytes_per_launch': nbytes,
               'gbs': nbytes / (med * 1e-3) / 1e9, 'frac': nbytes / (med * 1e-3) / 1e9 / PEAK_GBS}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--forms', default='fp32,fp16')
    ap.add_argument('--dims', default='512,256,256')
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--sweep', action='store_true')
    ap.add_argument('--shapes', default='', help="launch shapes to time, e.g. '1,2 2,10' (cpt,remap pairs)")
    ap.add_argument('--copy-gb', type=float, default=4.0)
    ap.add_argument('--particles', type=int, default=0)
    ap.add_argument('--fused', default='fp32', help='forms timed as fused step + advection with --particles')
    args = ap.parse_args(argv)
    dims = tuple(int(v) for v in args.dims.split(','))
    if args.copy_gb > 0:
        print(json.dumps(copy_bandwidth(args.copy_gb)), flush=True)
    mask = c2_mask(dims)
    for form in args.forms.split(','):
        # without --sweep: the shape tuned for this size class, the one sessions and bench.py run
        n = dims[0] * dims[1] * dims[2]
        shapes = ([(1, 0), (1, 1), (1, 2), (1, 3), (2, 0), (2, 2), (4, 0)] if args.sweep
                  else [autotune.shape_for(form, n)])
        if args.shapes:
            shapes = [tuple(int(v) for v in sh.split(',')) for sh in args.shapes.split()]
        t = None
        for shape in shapes:
            np_ = args.particles if form in args.fused.split(',') else 0
            t, r = step_timing(form, dims, mask, args.steps, tunnel=t, particles=np_, shape=shape)
            print(json.dumps(r), flush=True)
        del t
        torch.cuda.empty_cache()
    return 0


if __name__ == '__main__':
    sys.exit(main())
