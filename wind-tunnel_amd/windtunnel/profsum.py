"""Summaries of rocprofv3 output (CSV form) for the measurement records.

    python -m windtunnel.profsum stats <rocprof out dir> [--top 15]
        per-kernel calls / average / total time from *_kernel_stats.csv
    python -m windtunnel.profsum traffic <fetch dir> <write dir> --copy-bytes B [--json-dir D]
                                         [--copy-dirs FETCH_DIR WRITE_DIR] [--cells N] [--no-mac]
        HBM bytes per launch of each kernel from FETCH_SIZE and WRITE_SIZE
        (*_counter_collection.csv, KiB per dispatch), each scaled by what the
        same counter reads for wt_stream_copy of B bytes (reads B, writes B),
        from the same runs or, with --copy-dirs, from other runs on the box;
        batched steps (k_step_batch) are priced over --cells cells, without the
        macroscopic store with --no-mac

Prints markdown; with --json-dir also writes step_<form>_<dims>_traffic.json
(k_step) and step_adv_<form>_<dims>_traffic.json (k_step_adv, the step fused
with the advection of --particles particles) files that bench.py reports as
roofline.traffic.
"""
import argparse
import collections
import csv
import glob
import json
import os
import re
import sys

BYTES_PER_CELL = {0: 169, 1: 85, 2: 85, 3: 321}
MAC_BYTES = {0: 16, 1: 8, 2: 8, 3: 16}
BYTES_PER_PARTICLE = 40   # id 4, position 16 read + 16 written, speed 4 written
FORM_NAMES = {0: 'fp32', 1: 'fp16', 2: 'bf16', 3: 'fp64'}


def _rows(d, suffix):
    files = sorted(glob.glob(os.path.join(d, '**', '*' + suffix), recursive=True))
    if not files:
        raise SystemExit(f'no *{suffix} under {d}')
    out = []
    for f in files:
        with open(f, newline='') as fh:
            out += list(csv.DictReader(fh))
    return out


def short(name):
    name = re.sub(r'\(.*', '', name)
    return name.replace('void ', '').replace('wt::', '')


def stats(d, top=15):
    rows = _rows(d, 'kernel_stats.csv')
    rows.sort(key=lambda r: -float(r['TotalDurationNs']))
    lines = ['| kernel | calls | avg (us) | total (ms) | % |', '|---|---:|---:|---:|---:|']
    for r in rows[:top]:
        lines.append(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.2f} |")
    return '\n'.join(lines)


def _per_kernel(rows, counter):
    acc = collections.defaultdict(list)
    for r in rows:
        if r.get('Counter_Name') != counter:
            continue
        acc[(r['Kernel_Name'], int(r.get('Grid_Size', 0) or 0))].append(float(r['Counter_Value']))
    return acc


def traffic(fetch_dir, write_dir, copy_bytes, json_dir=None, cells=None, particles=0, copy_dirs=None,
            no_mac=False):
    fr = _per_kernel(_rows(fetch_dir, 'counter_collection.csv'), 'FETCH_SIZE')
    wr = _per_kernel(_rows(write_dir, 'counter_collection.csv'), 'WRITE_SIZE')
    cfr, cwr = fr, wr
    if copy_dirs:
        cfr = _per_kernel(_rows(copy_dirs[0], 'counter_collection.csv'), 'FETCH_SIZE')
        cwr = _per_kernel(_rows(copy_dirs[1], 'counter_collection.csv'), 'WRITE_SIZE')
    cf = [v for (k, _), vs in cfr.items() if 'k_stream_copy' in k for v in vs]
    cw = [v for (k, _), vs in cwr.items() if 'k_stream_copy' in k for v in vs]
    if not cf or not cw:
        raise SystemExit('no k_stream_copy dispatch in the counter runs')
    # median of the copy dispatches: KiB counted for copy_bytes read / written
    cf.sort(); cw.sort()
    f_scale = copy_bytes / cf[len(cf) // 2]
    w_scale = copy_bytes / cw[len(cw) // 2]
    lines = [f'Scaling: wt_stream_copy of {copy_bytes} B reads FETCH_SIZE {cf[len(cf) // 2]:.0f} and '
             f'WRITE_SIZE {cw[len(cw) // 2]:.0f} -> {f_scale:.2f} / {w_scale:.2f} bytes per counted unit', '',
             '| kernel | grid | dispatches | fetch (MB) | write (MB) | total (MB) | algorithmic (MB) | ratio |',
             '|---|---:|---:|---:|---:|---:|---:|---:|']
    out = {}
    for (name, grid), vs in sorted(fr.items()):
        if 'k_step' not in name:
            continue
        ws = wr.get((name, grid))
        if not ws:
            continue
        fb = sum(vs) / len(vs) * f_scale
        wb = sum(ws) / len(ws) * w_scale
        mb = re.search(r'k_step_batch<(\d)', name)
        if mb:
            form = int(mb.group(1))
            algo = (BYTES_PER_CELL[form] - (MAC_BYTES[form] if no_mac else 0)) * cells
            lines.append(f'| {short(name)} | {grid} | {len(vs)} | {fb / 1e6:.1f} | {wb / 1e6:.1f} | '
                         f'{(fb + wb) / 1e6:.1f} | {algo / 1e6:.1f} | {(fb + wb) / algo:.4f} |')
            continue
        m = re.search(r'k_step(_adv)?<(\d)', name)
        if not m:
            continue
        adv, form = bool(m.group(1)), int(m.group(2))
        algo = BYTES_PER_CELL[form] * cells + (BYTES_PER_PARTICLE * particles if adv else 0)
        lines.append(f'| {short(name)} | {grid} | {len(vs)} | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} | '
                     f'{algo / 1e6:.1f} | {(fb + wb) / algo:.4f} |')
        key = ('step_adv_' if adv else 'step_') + FORM_NAMES[form]
        out[key] = {'kernel': short(name), 'grid': grid, 'cells': cells, 'particles': particles if adv else 0,
                    'fetch_bytes': fb, 'write_bytes': wb, 'hbm_bytes_per_launch': fb + wb,
                    'bytes_per_cell': (fb + wb) / cells, 'algorithmic_bytes_per_launch': algo,
                    'algorithmic_bytes_per_cell': BYTES_PER_CELL[form],
                    'algorithmic_bytes_per_particle': BYTES_PER_PARTICLE if adv else 0,
                    'scale_fetch': f_scale, 'scale_write': w_scale}
    if json_dir:
        os.makedirs(json_dir, exist_ok=True)
        for key, v in out.items():
            with open(os.path.join(json_dir, f'{key}_512x256x256_traffic.json'), 'w') as fh:
                json.dump(v, fh, indent=1)
    return '\n'.join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    a = sub.add_parser('stats')
    a.add_argument('dir')
    a.add_argument('--top', type=int, default=15)
    b = sub.add_parser('traffic')
    b.add_argument('fetch_dir')
    b.add_argument('write_dir')
    b.add_argument('--copy-bytes', type=int, required=True)
    b.add_argument('--json-dir', default=None)
    b.add_argument('--cells', type=int, default=512 * 256 * 256)
    b.add_argument('--particles', type=int, default=0)
    b.add_argument('--copy-dirs', nargs=2, default=None, metavar=('FETCH_DIR', 'WRITE_DIR'))
    b.add_argument('--no-mac', action='store_true')
    args = ap.parse_args(argv)
    if args.cmd == 'stats':
        print(stats(args.dir, args.top))
    else:
        print(traffic(args.fetch_dir, args.write_dir, args.copy_bytes, args.json_dir, args.cells, args.particles,
                      args.copy_dirs, args.no_mac))
    return 0


if __name__ == '__main__':
    sys.exit(main())
