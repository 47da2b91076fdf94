"""The reference's numerical guards driven on the GPU, bit for bit against the
oracle in all four storage forms:

* the velocity clamp |u| <= 0.35 (lbm.py:130-135): a tunnel initialised at
  uin = 0.4 flows faster than the clamp in every interior cell;
* the face density clamp rho in [0.9, 1.1] (lbm.py:88): blocks of fluid at
  rho 1.3 and 0.8 next to the y and z walls, whose density the neighbouring face
  cells read in the following step.

The oracle counts how often each guard fired (oracle.clamp_hits), so the tests
assert that the guards really ran, not only that both sides agree."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FORMS = [('fp32', O.FP32), ('fp16', O.FP16), ('bf16', O.BF16), ('fp64', O.FP64)]
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C = O.C


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def _run(form, oform, dims, mask, uin, tau, smag, steps, raw=None):
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS
    from windtunnel.tunnel import Tunnel
    t = Tunnel(dims, form)
    t.init(uin)
    t.set_solid(mask)
    o = O.OracleTunnel(dims, oform)
    o.init(uin)
    o.set_solid(mask)
    if raw is not None:
        t.load_raw_populations(raw)
        o.f[0][:] = raw.reshape(-1)
        o.f[1][:] = raw.reshape(-1)
    O.clamp_hits()
    t.step(steps, uin, tau, smag, flags=WT_FLAG_FORCES | WT_FLAG_MASS)
    o.step(steps, uin, tau, smag)
    hits = O.clamp_hits()
    torch.cuda.synchronize()
    assert _same(t.raw_populations().cpu().numpy(), o.raw()), form
    assert _same(t.mac.cpu().numpy(), o.device_mac()), form
    sums = t.sums.cpu().numpy()
    np.testing.assert_allclose(sums[:3], o.forces(), rtol=1e-12, atol=1e-12 * np.abs(o.forces()).max())
    np.testing.assert_allclose(sums[3], o.mass, rtol=1e-12)
    return t, o, hits


@pytest.mark.parametrize('form,oform', FORMS)
def test_velocity_clamp_bit_exact(gpu, form, oform):
    from windtunnel import geometry
    dims = (64, 24, 26)
    mask, _ = geometry.sphere(dims, (20, 12, 13), 8)
    t, o, (vel, den) = _run(form, oform, dims, mask, 0.4, 0.52, 0.17, 3)
    nx, ny, nz = dims
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    # every interior fluid cell clamps in the first step, and many in the next two
    assert vel > interior, (vel, interior)
    m = o.mac.reshape(nz, ny, nx, 4)[1:-1, 1:-1, 1:-1].astype(np.float64)
    speed = np.linalg.norm(m[..., :3], axis=-1)            # faces keep the inlet's 0.4
    assert np.isclose(speed.max(), 0.35, rtol=1e-6)


def _dense_blocks(form, dims, uin):
    """Equilibrium populations at uin, with rho scaled to 1.3 in a block touching
    the wall y = ny - 2 and to 0.8 in a block next to z = 1, as storage elements."""
    nx, ny, nz = dims
    n = nx * ny * nz
    u = np.array([uin, 0.0, 0.0])
    cu = C @ u
    feq = W * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    rho = np.ones((nz, ny, nx))
    rho[4:12, ny - 4:ny - 1, 10:30] = 1.3
    rho[1:4, 3:10, 30:50] = 0.8
    f = feq[:, None] * rho.reshape(1, n)
    if form == 'fp32':
        return f.astype(np.float32)
    if form == 'fp64':
        return f
    g = f - W[:, None]
    if form == 'fp16':
        return g.astype(np.float16).view(np.uint16)
    b = g.astype(np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7fff + ((b >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.mark.parametrize('form,oform', FORMS)
def test_face_density_clamp_bit_exact(gpu, form, oform):
    dims = (64, 20, 18)
    mask = np.zeros(dims[::-1], np.uint8).reshape(-1)
    raw = _dense_blocks(form, dims, 0.05)
    t, o, (vel, den) = _run(form, oform, dims, mask, 0.05, 0.52, 0.17, 3, raw=raw)
    assert vel == 0
    assert den >= 400, den          # the face cells beside both blocks, steps 2 and 3
    nx, ny, nz = dims
    m = o.mac.reshape(nz, ny, nx, 4).astype(np.float64)
    assert np.isclose(m[:, ny - 1, 1:nx - 1, 3].max(), 1.1, rtol=1e-6)    # clamped from above
    assert np.isclose(m[0, 1:ny - 1, 1:nx - 1, 3].min(), 0.9, rtol=1e-6)  # clamped from below
