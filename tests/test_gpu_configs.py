"""The three multi-tunnel and precision workloads at their stated sizes.

* Precision study (384x192x192 = 14.2M cells, sphere D=48 at (120, 96, 96),
  uin 0.05, tau 0.52, Cs 0.17, no ground): two steps of every storage form bit
  for bit against the oracle at full size, and the velocity / pressure error of
  FP32, FP16 and BF16 against the FP64 form at step 200 within stated bounds.
* Batched tunnels (48 tunnels of uneven size drawn from {64x32x32, 96x48x48,
  128x64x64, 160x80x80}, sphere or ellipsoid, uin U(0.03, 0.08), tau
  U(0.51, 0.60)): three steps of the whole batch in one launch per
  step, every tunnel bit for bit against the oracle (populations, macroscopic
  field, FP64 force and mass), with the tuned launch shapes and with a
  different cells-per-thread per size class inside the one launch.
* Serving mix (Poisson sessions, 0.5M-33M cells log-uniform, half FP32 half
  FP16, memory-capped admission): sessions from that distribution, including
  its largest, with their step counts capped; each session's per-step forces
  are bit-identical alone and in the mix.
Reference anchor: Solver.step's per-tunnel launch loop, lbm.py:442-447."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FORMS = [('fp32', O.FP32), ('fp16', O.FP16), ('bf16', O.BF16), ('fp64', O.FP64)]
C3_DIMS = (384, 192, 192)
C3_UIN, C3_TAU, C3_SMAG = 0.05, 0.52, 0.17


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


@pytest.fixture(scope='module')
def c3_sphere():
    from windtunnel import geometry
    mask, counts = geometry.sphere(C3_DIMS, (120, 96, 96), 48)
    return mask, counts


@pytest.mark.parametrize('form,oform', FORMS)
def test_config3_two_steps_bit_exact_at_full_size(gpu, c3_sphere, form, oform):
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS
    from windtunnel.tunnel import Tunnel
    mask, _ = c3_sphere
    t = Tunnel(C3_DIMS, form)
    t.init(C3_UIN)
    t.set_solid(mask)
    t.step(2, C3_UIN, C3_TAU, C3_SMAG, flags=WT_FLAG_FORCES | WT_FLAG_MASS)
    f, mac, sums = t.raw_populations().cpu().numpy(), t.mac.cpu().numpy(), t.sums.cpu().numpy()
    del t
    torch.cuda.empty_cache()
    O.set_threads(16)
    o = O.OracleTunnel(C3_DIMS, oform)
    o.init(C3_UIN)
    o.set_solid(mask)
    o.step(2, C3_UIN, C3_TAU, C3_SMAG)
    assert _same(f, o.raw()), form
    assert _same(mac, o.device_mac()), form
    np.testing.assert_allclose(sums[:3], o.forces(), rtol=1e-12, atol=1e-12 * np.abs(o.forces()).max())
    np.testing.assert_allclose(sums[3], o.mass, rtol=1e-12)


# relative L2 error of the velocity field against the FP64 form at step 200, and
# max |dCp| over fluid cells (Cp = (rho - 1) / (1.5 uin^2): one float32 ulp of rho
# is 3.2e-5 in Cp); measured FP32 6.3e-6 / 1.0e-3, FP16 3.7e-4, BF16 3.2e-3
# (velocity): the bounds leave a factor of ~5
C3_BOUNDS = {'fp32': (3e-5, 5e-3), 'fp16': (2e-3, 0.1), 'bf16': (2e-2, 1.0)}


def test_config3_field_error_against_fp64_at_step_200(gpu, c3_sphere):
    from windtunnel.tunnel import Tunnel
    mask, _ = c3_sphere
    fluid = torch.from_numpy(mask == 0).cuda()
    fields = {}
    for form in ('fp64', 'fp32', 'fp16', 'bf16'):
        t = Tunnel(C3_DIMS, form)
        t.init(C3_UIN)
        t.set_solid(mask)
        t.step(200, C3_UIN, C3_TAU, C3_SMAG)
        m = t.macroscopic()[fluid].double()
        fields[form] = (m[:, :3].clone(), (m[:, 3] - 1.0) / (1.5 * C3_UIN * C3_UIN))
        del t, m
        torch.cuda.empty_cache()
    ref_u, ref_p = fields['fp64']
    err = {}
    for form in C3_BOUNDS:
        u, p = fields[form]
        assert torch.isfinite(u).all() and torch.isfinite(p).all(), form
        err[form] = (float((u - ref_u).norm() / ref_u.norm()), float((p - ref_p).abs().max()))
    print('config-3 errors against FP64 at step 200 (velocity L2 rel, max |dCp|):', err, flush=True)
    for form, (bu, bp) in C3_BOUNDS.items():
        assert err[form][0] < bu and err[form][1] < bp, err


# ---- batched tunnels ----------------------------------------------------------------------
def _c4_compare(form, oform, shapes=None, steps=3, persistent=False):
    from windtunnel import geometry
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS
    from windtunnel.batch import Batch
    from windtunnel.workloads import build_config4, config4_specs
    specs = config4_specs()
    assert len(specs) == 48 and len({s['dims'] for s in specs}) == 4
    tunnels, params = build_config4(form)
    if shapes:
        for t in tunnels:
            t.shape = shapes[(t.nx, t.ny, t.nz)]
    b = Batch(tunnels, params, own_shapes=bool(shapes), persistent=persistent)
    flags = WT_FLAG_FORCES | WT_FLAG_MASS
    for _ in range(steps):
        b.step(flags)
    torch.cuda.synchronize()
    sums = b.last_sums().cpu().numpy()
    O.set_threads(16)
    for k, (s, t) in enumerate(zip(specs, tunnels)):
        o = O.OracleTunnel(s['dims'], oform)
        o.init(s['uin'])
        o.set_solid(geometry.ellipsoids(s['ells'], s['dims'])[0])
        o.step(steps, s['uin'], s['tau'], s['smag'])
        assert _same(t.raw_populations().cpu().numpy(), o.raw()), (form, k, s['dims'])
        assert _same(t.mac.cpu().numpy(), o.device_mac()), (form, k)
        np.testing.assert_allclose(sums[k, :3], o.forces(), rtol=1e-12, atol=1e-12 * np.abs(o.forces()).max())
        np.testing.assert_allclose(sums[k, 3], o.mass, rtol=1e-12)


@pytest.mark.parametrize('persistent', [False, True])
def test_config4_batch_bit_exact_with_oracle(gpu, persistent):
    """The 48 tunnels (92.7k work items) by one workgroup per item, and by the
    persistent kernel, whose workgroups at full occupancy take ~60 items each."""
    _c4_compare('fp32', O.FP32, persistent=persistent)


MIXED = {(64, 32, 32): (1, 0), (96, 48, 48): (2, 0), (128, 64, 64): (4, 2), (160, 80, 80): (2, 6)}


@pytest.mark.parametrize('form,oform,persistent', [('fp32', O.FP32, False), ('fp16', O.FP16, False),
                                                   ('fp16', O.FP16, True)])
def test_config4_mixed_launch_shapes_in_one_launch(gpu, form, oform, persistent):
    _c4_compare(form, oform, shapes=MIXED, steps=2, persistent=persistent)


# ---- serving mix ----------------------------------------------------------------------------
def _c5_requests(steps=12):
    from windtunnel.serving import mixed_workload
    reqs = sorted(mixed_workload(seed=0, duration=600.0, rate=2.0), key=lambda r: int(np.prod(r.dims)))
    n = len(reqs)
    pick = [reqs[i] for i in (0, n // 5, 2 * n // 5, 3 * n // 5, 4 * n // 5, n - 2, n - 1)]
    # arrive in an order that makes the small ones wait behind the big ones
    pick.sort(key=lambda r: -int(np.prod(r.dims)))
    return [dataclasses.replace(r, sid=k, arrival=0.002 * k, steps=steps, slice_every=5) for k, r in enumerate(pick)]


def _alone(req):
    from windtunnel import geometry
    from windtunnel._lib import WT_FLAG_FORCES
    from windtunnel.tunnel import Tunnel
    mask, _ = geometry.ellipsoids(req.ellipsoids, req.dims, ground=req.ground)
    t = Tunnel(req.dims, req.form)
    t.init(req.uin)
    t.set_solid(mask)
    out = []
    for _ in range(req.steps):
        t.step(1, req.uin, req.tau, req.smag, flags=WT_FLAG_FORCES)
        out.append(t.sums.cpu().numpy()[:3].copy())
    del t
    torch.cuda.empty_cache()
    return np.array(out)


@pytest.mark.parametrize('graphs', [False, True])
def test_config5_sessions_identical_alone_and_in_the_mix(gpu, graphs):
    from windtunnel.serving import Server
    from windtunnel.tunnel import state_bytes
    reqs = _c5_requests()
    cells = [int(np.prod(r.dims)) for r in reqs]
    assert min(cells) < 1.0e6 and max(cells) > 25e6 and len({r.form for r in reqs}) == 2
    foot = sorted(state_bytes(r.dims, r.form) + 4 * int(np.prod(r.dims)) for r in reqs)
    srv = Server(budget_bytes=foot[-1] + foot[-2], use_graphs=graphs)   # the two largest fit together
    st = srv.run(reqs)
    assert st['sessions_done'] == len(reqs)
    assert max(srv.queue_delays) > 0.0                                  # someone waited for memory
    assert st['cell_steps'] == sum(r.steps * c for r, c in zip(reqs, cells))
    for r in reqs:
        lv = srv.finished[r.sid]
        assert lv.done_steps == r.steps
        assert np.array_equal(np.array(lv.forces).view(np.uint64), _alone(r).view(np.uint64)), (r.sid, r.dims)
        assert lv.slices == r.steps // r.slice_every
