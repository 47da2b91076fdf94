"""Launch-shape autotuner for the fused step kernel.

Shapes: cells per thread (1, 2 or 4 consecutive 256-cell slots per
workgroup) x tile order (bit 0: XCD-contiguous; bit 1: sweep direction
alternating every step). Results do not depend on the
shape (the force/mass slots are reduced identically), only the speed does.
For each storage form and size class (classes follow the cache policy the
library uses for the tunnel's size) the tuner times every shape on a
representative tunnel and stores the fastest in tuning.json beside this file;
tunnels look their shape up by size class when they are created.

Batches are tuned apart (tune_batch), on a batch that spans the size classes:
first one cells-per-thread value for every tunnel of a batch of that storage
form, then, from there, a value per size class (each class in turn tries every
value while the others keep theirs, kept if it is faster), which a batch gives
each tunnel by its class. The batched kernel is compiled per largest
cells-per-thread of its batch, and its one-cell variant holds more waves per
SIMD, so a batch does not take its members' single-tunnel shapes (an FP16 batch
of 60 serving-mix tunnels measured 7 % faster at one cell per thread than with
the L class's two).

    python -m windtunnel.autotune [--forms fp32,fp16,bf16,fp64] [--steps 40] [--batch-only]
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
# the batch tune_batch times: tunnels as batches hold them, 16 of the serving mix's
# sizes (log-uniform 0.5M-33M cells, 2m x m x m) and 16 of the 48-tunnel batch's
BATCH_SERVING, BATCH_SMALL = 16, 16
BATCH_SMALL_DIMS = [(64, 32, 32), (96, 48, 48), (128, 64, 64), (160, 80, 80)]
BATCH_CPTS = (1, 2, 4)

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


def _load_file():
    global _table
    if _table is None:
        try:
            with open(TUNING_PATH) as fh:
                _table = json.load(fh)
        except (OSError, ValueError):
            _table = {}
    return _table


def load():
    """The single-tunnel shapes: form -> size class -> (cells per thread, tile order)."""
    return _load_file().get('shapes', {})


def shape_for(form_name, cells):
    t = load().get(form_name, {})
    v = t.get(size_class(cells, form_name))
    return tuple(v) if v else DEFAULT


def batch_cpt(form_name, cells=None):
    """Cells per thread of a tunnel of this many cells in a batch of this storage
    form: the value tuned for its size class, else the form's batch value."""
    t = _load_file()
    per_class = t.get('batch_cpt_class', {}).get(form_name)
    if cells is not None and per_class:
        v = per_class.get(size_class(cells, form_name))
        if v:
            return int(v)
    return int(t.get('batch_cpt', {}).get(form_name, 1))


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


def _class_mask(dims):
    from . import geometry
    ells = geometry.car_body(dims, 7, count=8,
                             box=((0.2 * dims[0], 0.5 * dims[0]), (1, 0.3 * dims[1]),
                                  (0.35 * dims[2], 0.65 * dims[2])),
                             axes=(0.03 * dims[1], 0.12 * dims[1]))
    return geometry.ellipsoids(ells, dims, ground=True)[0]


def batch_dims(seed=3):
    """Grid sizes of the tuning batch (BATCH_SERVING + BATCH_SMALL tunnels)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(BATCH_SERVING):
        cells = float(np.exp(rng.uniform(np.log(0.5e6), np.log(33e6))))
        m = max(16, int(round((cells / 2) ** (1 / 3) / 8)) * 8)
        out.append((2 * m, m, m))
    out += [BATCH_SMALL_DIMS[int(k)] for k in rng.integers(len(BATCH_SMALL_DIMS), size=BATCH_SMALL)]
    return out


def tune_batch(forms, steps=20):
    """Tune the cells per thread of batched tunnels on one batch per storage form
    (batch_dims(): tunnels of the sizes batches hold, forces every step, no
    macroscopic stores): a uniform value first, then one per size class by
    coordinate descent from it. Returns {'batch_cpt': {form: uniform best},
    'batch_cpt_class': {form: {class: cpt}}, 'batch_timings': {form: {...: ms}}}."""
    import numpy as np
    import torch
    from .batch import Batch
    from .tunnel import Tunnel
    out = {'batch_cpt': {}, 'batch_cpt_class': {}, 'batch_timings': {}}
    for form in forms:
        tunnels = []
        for dims in batch_dims():
            t = Tunnel(dims, form)
            t.init(0.05)
            t.set_solid(_class_mask(dims))
            tunnels.append(t)
        params = [(0.05, 0.52, 0.16)] * len(tunnels)
        classes = [size_class(t.n, form) for t in tunnels]

        def time_of(cpts):
            # cpts: cells per thread per tunnel; median of two runs
            runs = []
            for _ in range(2):
                for t, c in zip(tunnels, cpts):
                    t.shape = (c, t.shape[1])
                b = Batch(tunnels, params, own_shapes=True)
                b.set_gates([False] * len(tunnels))
                for _ in range(3):
                    b.step(1 | 4)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(steps):
                    b.step(1 | 4)
                e1.record()
                torch.cuda.synchronize()
                runs.append(e0.elapsed_time(e1) / steps)
                del b
            return float(np.median(runs))

        ms = {}
        for cpt in BATCH_CPTS:
            ms[str(cpt)] = time_of([cpt] * len(tunnels))
        best = min(BATCH_CPTS, key=lambda c: ms[str(c)])
        per_class = {c: best for c in sorted(set(classes))}
        cur = ms[str(best)]
        for cl in sorted(set(classes)):
            for cpt in BATCH_CPTS:
                if cpt == per_class[cl]:
                    continue
                trial = dict(per_class, **{cl: cpt})
                key = ','.join(f'{k}{v}' for k, v in sorted(trial.items()))
                ms[key] = time_of([trial[c] for c in classes])
                # keep only a clear gain: a class whose value raises the batch's largest
                # cells per thread moves every tunnel to the slower kernel variant, which the
                # tuning batch can hide within ~0.5 % (FP64: 0.56 % faster here, 1.6 % slower
                # on the 48-tunnel batch)
                if ms[key] < cur * 0.99:
                    per_class, cur = trial, ms[key]
        out['batch_cpt'][form] = best
        out['batch_cpt_class'][form] = per_class
        out['batch_timings'][form] = ms
        print(json.dumps({'form': form, 'batch_tunnels': len(tunnels), 'best_cpt': best, 'per_class': per_class,
                          'ms': ms}), flush=True)
        del tunnels
        torch.cuda.empty_cache()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--forms', default='fp32,fp16,bf16,fp64')
    ap.add_argument('--steps', type=int, default=40)
    ap.add_argument('--out', default=TUNING_PATH)
    ap.add_argument('--batch-only', action='store_true',
                    help='re-tune only the batch entry, keeping the single-tunnel shapes of --out')
    args = ap.parse_args(argv)
    forms = args.forms.split(',')
    if args.batch_only:
        with open(args.out) as fh:
            res = json.load(fh)
    else:
        res = tune(forms, args.steps)
    b = tune_batch(forms)
    for key in ('batch_cpt', 'batch_cpt_class', 'batch_timings'):
        res.setdefault(key, {}).update(b[key])
    with open(args.out, 'w') as fh:
        json.dump(res, fh, indent=1)
    return 0


if __name__ == '__main__':
    sys.exit(main())
