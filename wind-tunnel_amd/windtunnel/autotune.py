"""Launch-shape autotuner for the fused step kernel.

Shapes: cells per thread (1, 2 or 4 consecutive 256-cell slots per
workgroup) x tile order (bit 0: XCD-contiguous; bit 1: sweep direction
alternating every step). Results do not depend on the
shape (the force/mass slots are reduced identically), only the speed does.
For each storage form and size class (classes follow the cache policy the
library uses for the tunnel's size) the tuner times every shape on a
representative tunnel and stores the fastest in tuning.json beside this file;
tunnels look their shape up by size class when they are created.

    python -m windtunnel.autotune [--forms fp32,fp16,bf16,fp64] [--steps 40]
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.environ.get('WT_TUNING_PATH') or os.path.join(HERE, 'tuning.json')

# Size classes follow the cache policy the library picks for a tunnel (the bytes
# of its two population buffers, wt_cache_policy), so that no class straddles a
# policy boundary: S and M are all-cached (policy 2; S below 200k cells), L the
# middle policy 0, XL the streaming policy 1. The tunnel each is tuned on:
CLASSES = [('S', (64, 32, 32)), ('M', (128, 64, 64)), ('L', (256, 128, 128)), ('XL', (512, 256, 256))]
SMALL_CELLS = 200_000
ELEM_BYTES = {'fp32': 4, 'fp16': 2, 'bf16': 2, 'fp64': 8}
SHAPES = [(1, 0), (1, 1), (1, 2), (1, 3), (2, 0), (2, 1), (2, 2), (4, 0), (4, 1)]
DEFAULT = (1, 0)

_table = None


def cache_policy(form_name, cells):
    """The library's single-tunnel cache policy for this many cells (the rule of
    wt_cache_policy: 1 from 2 GiB of populations, 2 within 224 MiB, else 0)."""
    b = 2 * 19 * int(cells) * ELEM_BYTES[form_name]
    return 1 if b >= 1 << 31 else (2 if b <= 224 << 20 else 0)


def size_class(cells, form_name='fp32'):
    pol = cache_policy(form_name, cells)
    if pol == 1:
        return 'XL'
    if pol == 0:
        return 'L'
    return 'S' if cells < SMALL_CELLS else 'M'


def load():
    global _table
    if _table is None:
        try:
            with open(TUNING_PATH) as fh:
                _table = json.load(fh).get('shapes', {})
        except (OSError, ValueError):
            _table = {}
    return _table


def shape_for(form_name, cells):
    t = load().get(form_name, {})
    v = t.get(size_class(cells, form_name))
    return tuple(v) if v else DEFAULT


def tune(forms, steps=40):
    import torch
    from . import geometry
    from .roofline import step_timing
    from .tunnel import Tunnel
    out = {'arch': torch.cuda.get_device_properties(0).gcnArchName, 'shapes': {}, 'timings': {}}
    for form in forms:
        out['shapes'][form] = {}
        out['timings'][form] = {}
        for name, dims in CLASSES:
            assert size_class(dims[0] * dims[1] * dims[2], form) == name, (form, name, dims)
            ells = geometry.car_body(dims, 7, count=8,
                                     box=((0.2 * dims[0], 0.5 * dims[0]), (1, 0.3 * dims[1]),
                                          (0.35 * dims[2], 0.65 * dims[2])),
                                     axes=(0.03 * dims[1], 0.12 * dims[1]))
            mask = geometry.ellipsoids(ells, dims, ground=True)[0]
            t = Tunnel(dims, form)
            t.init(0.05)
            t.set_solid(mask)
            best, res = None, {}
            for shape in SHAPES:
                _, r = step_timing(form, dims, mask, steps=steps, warmup=5, tunnel=t, shape=shape)
                res['%d,%d' % shape] = r['ms_median']
                if best is None or r['ms_median'] < res['%d,%d' % best]:
                    best = shape
            out['shapes'][form][name] = list(best)
            out['timings'][form][name] = res
            print(json.dumps({'form': form, 'class': name, 'dims': dims, 'best': best, 'ms': res}), flush=True)
            del t
            torch.cuda.empty_cache()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--forms', default='fp32,fp16,bf16,fp64')
    ap.add_argument('--steps', type=int, default=40)
    ap.add_argument('--out', default=TUNING_PATH)
    args = ap.parse_args(argv)
    res = tune(args.forms.split(','), args.steps)
    with open(args.out, 'w') as fh:
        json.dump(res, fh, indent=1)
    return 0


if __name__ == '__main__':
    sys.exit(main())
