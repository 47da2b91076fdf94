"""Benchmark workloads beyond the single large tunnel.

config4: 48 independent tunnels of uneven size drawn by seed from
{64x32x32, 96x48x48, 128x64x64, 160x80x80}, each with a seeded sphere or
ellipsoid obstacle, inlet speed U(0.03, 0.08), tau U(0.51, 0.60), FP32; all of
them advance in ONE batched launch per step, with forces every step fetched
in one copy. Nothing in the workload reads the macroscopic field, so its steps
do not store it unless --store-mac asks for it (as the reference's step does).

    python -m windtunnel.workloads [--steps 1000] [--warmup 20] [--form fp32] [--no-graphs] [--store-mac]
"""
import time

import numpy as np
import torch

from . import geometry
from ._lib import WT_FLAG_FORCES, WT_FLAG_NO_MAC
from .batch import Batch
from .tunnel import Tunnel

C4_SIZES = [(64, 32, 32), (96, 48, 48), (128, 64, 64), (160, 80, 80)]


def config4_specs(seed=4, ntunnels=48):
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(ntunnels):
        dims = C4_SIZES[int(rng.integers(len(C4_SIZES)))]
        nx, ny, nz = dims
        if rng.random() < 0.5:
            d = rng.uniform(0.2, 0.35) * ny
            ells = [[0.3 * nx, ny / 2, nz / 2, d / 2, d / 2, d / 2]]
        else:
            ells = [[0.3 * nx, ny / 2, nz / 2, rng.uniform(0.05, 0.15) * nx, rng.uniform(0.1, 0.25) * ny,
                     rng.uniform(0.1, 0.25) * nz]]
        specs.append({'dims': dims, 'ells': np.array(ells), 'uin': float(rng.uniform(0.03, 0.08)),
                      'tau': float(rng.uniform(0.51, 0.60)), 'smag': 0.16})
    return specs


def build_config4(form='fp32', seed=4, ntunnels=48):
    tunnels, params = [], []
    for s in config4_specs(seed, ntunnels):
        t = Tunnel(s['dims'], form)
        t.init(s['uin'])
        t.set_solid(geometry.ellipsoids(s['ells'], s['dims'])[0])
        tunnels.append(t)
        params.append((s['uin'], s['tau'], s['smag']))
    return tunnels, params


def run_config4(steps=1000, warmup=20, form='fp32', graphs=False, seed=4, store_mac=False, persistent=None):
    """The workload asks for forces every step and nothing that reads the
    macroscopic field, so by default no step stores it (WT_FLAG_NO_MAC, every gate
    closed); store_mac=True stores it every step as the reference's step does."""
    tunnels, params = build_config4(form, seed)
    b = Batch(tunnels, params, persistent=persistent)
    flags = WT_FLAG_FORCES
    if not store_mac:
        flags |= WT_FLAG_NO_MAC
        b.set_gates([False] * len(tunnels))
    if graphs:
        b.capture(flags)
    host = torch.zeros((2, len(tunnels), 4), dtype=torch.float64).pin_memory()

    def advance(k):
        if graphs:
            b.replay()
            return 2
        b.step(flags)
        return 1

    k = 0
    while k < warmup:
        k += advance(k)
    torch.cuda.synchronize()
    cells = b.cells
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    h0 = time.perf_counter()
    e0.record()
    done = 0
    while done < steps:
        done += advance(done)
        host.copy_(b.sums, non_blocking=True)      # forces of every tunnel, every step
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - h0
    return {'tunnels': len(tunnels), 'cells': cells, 'steps': done, 'graphs': graphs, 'store_mac': store_mac,
            'persistent': b.persistent,
            'wall_s': wall, 'gpu_ms': e0.elapsed_time(e1), 'mlups': cells * done / wall / 1e6,
            'ms_per_step': wall * 1e3 / done}


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description='config-4 batch: 48 uneven tunnels, one batched launch per step')
    ap.add_argument('--steps', type=int, default=1000)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--form', default='fp32')
    ap.add_argument('--no-graphs', action='store_true')
    ap.add_argument('--store-mac', action='store_true', help='store the macroscopic field every step')
    ap.add_argument('--persistent', type=int, choices=(0, 1), default=None,
                    help='1: the persistent batched kernel, 0: one workgroup per item (default: Batch\'s)')
    args = ap.parse_args(argv)
    print(json.dumps(run_config4(args.steps, args.warmup, args.form, graphs=not args.no_graphs,
                                 store_mac=args.store_mac, persistent=args.persistent)), flush=True)
    return 0


if __name__ == '__main__':
    import sys
    sys.exit(main())
