"""The benchmark workload at its full size: 512x256x256 = 33.5M cells, the
12-ellipsoid car body (seed 2025) on the moving belt, uin 0.05, tau 0.51,
Cs 0.16, the launch shape tuned for this size class.

* Two steps with each storage form (FP32, FP16 and BF16 deviations, FP64) match
  the oracle bit for bit (populations, mac); FP64 forces within rtol 1e-12
  (summation order).
* Properties that hold at any size, over the bench loop (2^20 particles, three
  slices and the force every 10 steps): fusing each step's advection into the
  next launch changes no bit of any output; another launch shape changes no
  bit; every particle stays finite and inside the tunnel."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIMS = (512, 256, 256)
UIN, TAU, SMAG = 0.05, 0.51, 0.16


@pytest.fixture(scope='module')
def body():
    from windtunnel import geometry
    ells = geometry.car_body(DIMS, 2025)
    mask, counts = geometry.ellipsoids(ells, DIMS, ground=True)
    return mask, counts


@pytest.mark.parametrize('form,oform', [('fp32', O.FP32), ('fp16', O.FP16), ('bf16', O.BF16), ('fp64', O.FP64)])
def test_two_steps_bit_exact_with_oracle(gpu, body, form, oform):
    from windtunnel._lib import WT_FLAG_FORCES
    from windtunnel.tunnel import Tunnel
    mask, _ = body
    t = Tunnel(DIMS, form)
    assert t.shape is not None              # the tuned shape of this size class
    t.init(UIN)
    t.set_solid(mask)
    t.step(2, UIN, TAU, SMAG, flags=WT_FLAG_FORCES)
    f = t.raw_populations().cpu().numpy()
    mac = t.mac.cpu().numpy()
    sums = t.sums.cpu().numpy()
    del t
    torch.cuda.empty_cache()
    O.set_threads(16)
    o = O.OracleTunnel(DIMS, oform)
    o.init(UIN)
    o.set_solid(mask)
    o.step(2, UIN, TAU, SMAG)
    raw = o.raw()
    assert f.dtype.itemsize == raw.dtype.itemsize
    assert np.array_equal(f.view(np.uint8), raw.view(np.uint8))
    omac = o.device_mac()     # float32 (FP32, FP64) or float16 with rho - 1 (FP16, BF16)
    assert mac.dtype == omac.dtype and np.array_equal(mac.reshape(-1).view(np.uint8), omac.reshape(-1).view(np.uint8))
    assert np.allclose(sums[:3], o.forces()[:3], rtol=1e-12, atol=0)


def _session(body, fuse, steps):
    from windtunnel.session import Session
    mask, counts = body
    nz = DIMS[2]
    s = Session(DIMS, mask, 'fp32', UIN, TAU, SMAG, particles=1 << 20,
                slices=[(2, nz // 2, 0), (2, nz // 2, 1), (2, nz // 2, 2)], frame_every=10, lref=100.0,
                seed=7, frontal=counts['frontal'], fuse=fuse)
    for _ in range(steps):
        s.step()
    return s


def test_fused_bench_loop_equals_separate_launches(gpu, body):
    a = _session(body, True, 120)
    pa, sa = a.particles_by_id()
    b = _session(body, False, 120)
    pb, sb = b.particles_by_id()
    torch.cuda.synchronize()
    assert a.T.cur == b.T.cur
    assert torch.equal(a.T.f[a.T.cur], b.T.f[b.T.cur])
    assert torch.equal(a.T.mac, b.T.mac) and torch.equal(a.T.sums, b.T.sums)
    assert pa.tobytes() == pb.tobytes() and sa.tobytes() == sb.tobytes()
    for x, y in zip(a.slice_host, b.slice_host):
        assert torch.equal(x, y)
    assert np.isfinite(pa).all() and (sa > 0).all()
    nx, ny, nz = DIMS
    assert (pa[:, 0] >= 0.5).all() and (pa[:, 0] <= nx - 3.0).all()
    assert (pa[:, 1] >= 0.5).all() and (pa[:, 1] <= ny - 1.5).all()
    assert (pa[:, 2] >= 0.5).all() and (pa[:, 2] <= nz - 1.5).all()


def test_launch_shape_changes_no_bit_at_full_size(gpu, body):
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS
    from windtunnel.tunnel import Tunnel
    mask, _ = body
    out = []
    for shape in ((1, 2), (2, 1)):
        t = Tunnel(DIMS, 'fp32')
        t.shape = shape
        t.init(UIN)
        t.set_solid(mask)
        t.step(7, UIN, TAU, SMAG, flags=WT_FLAG_FORCES | WT_FLAG_MASS)
        out.append(t)
    torch.cuda.synchronize()
    a, b = out
    assert torch.equal(a.f[a.cur], b.f[b.cur]) and torch.equal(a.mac, b.mac)
    assert torch.equal(a.sums, b.sums)
