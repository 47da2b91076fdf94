"""Precision study of the population storage forms.

Flow past a sphere (the 128x64x64 / D=16 case rescaled by `scale`; default 3:
384x192x192, D=48 at (120, 96, 96)), uin 0.05, tau 0.52, Cs 0.17, no ground.
Every form (FP32; FP16 and BF16 deviations from w_i with FP32 collision; FP64)
runs the same steps. Reported against the FP64 form:
  * velocity and pressure (Cp = (rho - 1) / (1.5 uin^2)) L2-relative and max
    errors over fluid cells at step `field_step`;
  * Cd, Cl = F / (0.5 uin^2 A_frontal), averaged over steps [avg_from, avg_to),
    forces every step, and their relative difference to FP64's.

    python -m windtunnel.precision [--scale 3] [--field-step 200] [--avg-from 1000] [--avg-to 3000]
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


def run(scale=3, field_step=200, avg_from=1000, avg_to=3000, uin=0.05, tau=0.52, smag=0.17, forms=FORMS):
    dims = (128 * scale, 64 * scale, 64 * scale)
    mask, counts = geometry.sphere(dims, (40 * scale, 32 * scale, 32 * scale), 16 * scale)
    fluid = torch.from_numpy(mask == 0).cuda()
    q = 0.5 * uin * uin * counts['frontal']
    res, fields = {}, {}
    for form in forms:
        t = Tunnel(dims, form)
        t.init(uin)
        t.set_solid(mask)
        fx, fy = [], []
        f_buf = torch.zeros((avg_to, 4), dtype=torch.float64, device='cuda')
        for k in range(1, avg_to + 1):
            want = k >= avg_from
            t.step(1, uin, tau, smag, flags=WT_FLAG_FORCES if want else 0)
            if want:
                f_buf[k - 1].copy_(t.sums)
            if k == field_step:
                m = t.mac[fluid].double()
                fields[form] = (m[:, :3].clone(), ((m[:, 3] - 1.0) / (1.5 * uin * uin)).clone())
        torch.cuda.synchronize()
        f = f_buf[avg_from - 1:avg_to].cpu().numpy()
        res[form] = {'cd': float(f[:, 0].mean() / q), 'cl': float(f[:, 1].mean() / q),
                     'cd_std': float(f[:, 0].std() / q), 'finite': bool(np.isfinite(f).all())}
        del t
        torch.cuda.empty_cache()
    ref_u, ref_p = fields['fp64']
    for form in forms:
        u, p = fields[form]
        du = (u - ref_u).norm(dim=1)
        res[form]['u_l2_rel'] = float(du.norm() / ref_u.norm(dim=1).norm())
        res[form]['u_max_abs_over_uin'] = float(du.max() / uin)
        res[form]['p_l2_rel'] = float((p - ref_p).norm() / ref_p.norm())
        res[form]['p_max_abs'] = float((p - ref_p).abs().max())
        for c in ('cd', 'cl'):
            res[form][c + '_rel_to_fp64'] = abs(res[form][c] - res['fp64'][c]) / max(abs(res['fp64'][c]), 1e-30)
    return {'dims': dims, 'sphere_diameter': 16 * scale, 'frontal': counts['frontal'], 'uin': uin, 'tau': tau,
            'smag': smag, 'field_step': field_step, 'avg_steps': [avg_from, avg_to], 'forms': res}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--scale', type=int, default=3)
    ap.add_argument('--field-step', type=int, default=200)
    ap.add_argument('--avg-from', type=int, default=1000)
    ap.add_argument('--avg-to', type=int, default=3000)
    ap.add_argument('--forms', default=','.join(FORMS))
    args = ap.parse_args(argv)
    forms = args.forms.split(',')
    if forms[0] != 'fp64':
        forms = ['fp64'] + [f for f in forms if f != 'fp64']
    print(json.dumps(run(args.scale, args.field_step, args.avg_from, args.avg_to, forms=forms)), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
