"""Precision of the population storage forms against the FP64 form.

Flow past a sphere: the 128x64x64 tunnel with a D=16 sphere at (40, 32, 32)
rescaled by `scale` (3: 384x192x192, D=48 at (120, 96, 96)), uin 0.05,
tau 0.52, Cs 0.17, no ground. Every form (FP32; FP16 and BF16 deviations from
w_i with FP32 collision; FP64) runs the same steps from the same start.

field_errors: velocity and pressure (Cp = (rho - 1) / (1.5 uin^2)) at step
`step` over fluid cells, against the FP64 form: relative L2 of the velocity
difference and max |dCp|.

coefficients: Cd, Cl = F / (0.5 uin^2 A_frontal) (web/js/main.js:336-339) with
the force every step, averaged over steps [avg_from, avg_to) -- a window that
starts after at least two flow-throughs (nx / uin steps each) -- with their
standard deviation over the window and the relative difference of the means to
the FP64 form's. At scale 1 (Re = uin D / nu = 120) the wake is steady and the
averages converge; at scale 3 (Re 360) it sheds and the window average is
reported with its spread.

    python -m windtunnel.precision fields [--scale 3] [--step 200]
    python -m windtunnel.precision coefficients [--scale 1] [--avg-from 6000] [--avg-to 10000]
"""
import argparse
import json
import sys

import numpy as np
import torch

from . import geometry
from ._lib import WT_FLAG_FORCES
from .tunnel import Tunnel

FORMS = ('fp64', 'fp32', 'fp16', 'bf16')
UIN, TAU, SMAG = 0.05, 0.52, 0.17


def sphere_case(scale):
    dims = (128 * scale, 64 * scale, 64 * scale)
    mask, counts = geometry.sphere(dims, (40 * scale, 32 * scale, 32 * scale), 16 * scale)
    return dims, mask, counts


def _tunnel(dims, mask, form):
    t = Tunnel(dims, form)
    t.init(UIN)
    t.set_solid(mask)
    return t


def field_errors(scale=3, step=200, forms=FORMS):
    """{form: {'u_l2_rel', 'u_max_abs_over_uin', 'p_l2_rel', 'p_max_abs'}} at `step`."""
    dims, mask, _ = sphere_case(scale)
    fluid = torch.from_numpy(mask == 0).cuda()
    fields = {}
    for form in ('fp64',) + tuple(f for f in forms if f != 'fp64'):
        t = _tunnel(dims, mask, form)
        t.step(step, UIN, TAU, SMAG)
        m = t.macroscopic()[fluid].double()
        fields[form] = (m[:, :3].clone(), (m[:, 3] - 1.0) / (1.5 * UIN * UIN))
        del t, m
        torch.cuda.empty_cache()
    ref_u, ref_p = fields['fp64']
    out = {}
    for form, (u, p) in fields.items():
        if form == 'fp64':
            continue
        du = (u - ref_u).norm(dim=1)
        out[form] = {'u_l2_rel': float(du.norm() / ref_u.norm(dim=1).norm()),
                     'u_max_abs_over_uin': float(du.max() / UIN),
                     'p_l2_rel': float((p - ref_p).norm() / ref_p.norm()),
                     'p_max_abs': float((p - ref_p).abs().max())}
    return {'dims': dims, 'step': step, 'forms': out}


def coefficients(scale=1, avg_from=6000, avg_to=10000, forms=FORMS):
    """Time-averaged Cd/Cl of every form over [avg_from, avg_to) and their
    relative difference to the FP64 form's."""
    dims, mask, counts = sphere_case(scale)
    q = 0.5 * UIN * UIN * counts['frontal']
    res = {}
    for form in ('fp64',) + tuple(f for f in forms if f != 'fp64'):
        t = _tunnel(dims, mask, form)
        if avg_from > 0:
            t.step(avg_from, UIN, TAU, SMAG)
        log = torch.zeros((avg_to - avg_from, 4), dtype=torch.float64, device='cuda')
        for k in range(avg_to - avg_from):
            t.step(1, UIN, TAU, SMAG, flags=WT_FLAG_FORCES)
            log[k].copy_(t.sums)
        f = log.cpu().numpy()
        del t, log
        torch.cuda.empty_cache()
        res[form] = {'cd': float(f[:, 0].mean() / q), 'cl': float(f[:, 1].mean() / q),
                     'cd_std': float(f[:, 0].std() / q), 'cl_std': float(f[:, 1].std() / q),
                     'finite': bool(np.isfinite(f).all())}
    cd64 = res['fp64']['cd']
    for form, r in res.items():
        r['cd_rel_to_fp64'] = abs(r['cd'] - cd64) / abs(cd64)
        r['cl_minus_fp64'] = r['cl'] - res['fp64']['cl']
    return {'dims': dims, 'sphere_diameter': 16 * scale, 'frontal': counts['frontal'], 'uin': UIN, 'tau': TAU,
            'smag': SMAG, 'reynolds': UIN * 16 * scale / ((TAU - 0.5) / 3), 'flow_through_steps': dims[0] / UIN,
            'avg_steps': [avg_from, avg_to], 'forms': res}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    a = sub.add_parser('fields')
    a.add_argument('--scale', type=int, default=3)
    a.add_argument('--step', type=int, default=200)
    b = sub.add_parser('coefficients')
    b.add_argument('--scale', type=int, default=1)
    b.add_argument('--avg-from', type=int, default=6000)
    b.add_argument('--avg-to', type=int, default=10000)
    for p in (a, b):
        p.add_argument('--forms', default=','.join(FORMS))
    args = ap.parse_args(argv)
    forms = tuple(args.forms.split(','))
    if args.cmd == 'fields':
        out = field_errors(args.scale, args.step, forms)
    else:
        out = coefficients(args.scale, args.avg_from, args.avg_to, forms)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
