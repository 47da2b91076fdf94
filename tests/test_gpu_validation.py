"""The Tunnel wrappers check every tensor they hand to a kernel as a raw
pointer: a buffer on the wrong device, of the wrong dtype or shape, not
contiguous or too small raises ValueError on the host, before any launch."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DIMS = (24, 12, 10)
BOX = (2.0, 6.0, 5.0, 3.0, 3.0)


@pytest.fixture
def tun(gpu):
    from windtunnel.tunnel import Tunnel
    t = Tunnel(DIMS, 'fp32')
    t.init(0.05)
    t.set_solid(np.zeros(DIMS[::-1], np.uint8))
    return t


def _parts(n=32, dev='cuda'):
    pos = torch.zeros((n, 4), dtype=torch.float32, device=dev)
    pos[:, 0], pos[:, 1], pos[:, 2] = 3.0, 6.0, 5.0
    return pos, torch.ones(n, dtype=torch.float32, device=dev), torch.arange(n, dtype=torch.int32, device=dev)


def test_particle_buffers_are_checked(tun):
    pos, spd, ids = _parts()
    bad = {
        'pos (n, 3)': (pos[:, :3].contiguous(), spd, ids),
        'pos on the host': (pos.cpu(), spd, ids),
        'pos float64': (pos.double(), spd, ids),
        'pos not contiguous': (torch.zeros((4, 32), device='cuda').t(), spd, ids),
        'spd too short': (pos, spd[:-1], ids),
        'spd float16': (pos, spd.half(), ids),
        'ids int64': (pos, spd, ids.long()),
        'ids too long': (pos, spd, torch.arange(33, dtype=torch.int32, device='cuda')),
    }
    for what, (p, s, i) in bad.items():
        with pytest.raises(ValueError):
            tun.advect_slots(p, s, i, 1.0, 0.05, 100.0, 1, BOX)
        if what.startswith('ids'):
            continue
        with pytest.raises(ValueError):
            tun.advect(p, s, 1.0, 0.05, 100.0, 1, BOX)
    tun.advect_slots(pos, spd, ids, 1.0, 0.05, 100.0, 1, BOX)     # the well-formed call still runs
    torch.cuda.synchronize()
    assert torch.isfinite(pos).all()


def test_gather_order_pack_and_fused_step_are_checked(tun):
    pos, spd, ids = _parts()
    order = torch.arange(32, dtype=torch.int32, device='cuda')
    out = _parts()
    with pytest.raises(ValueError):
        tun.particle_gather(pos, spd, ids, order.long(), *out)
    with pytest.raises(ValueError):
        tun.particle_gather(pos, spd, ids, order, *_parts(31))
    with pytest.raises(ValueError):
        tun.particle_order(pos, order[:16])
    q16 = torch.zeros(4 * 32, dtype=torch.int16, device='cuda')
    q8 = torch.zeros(32, dtype=torch.uint8, device='cuda')
    with pytest.raises(ValueError):
        tun.pack_particles_slots(pos, spd, ids, q16[:100], q8)
    with pytest.raises(ValueError):
        tun.pack_particles(pos, spd, q16, q8.to(torch.int8))
    other = torch.zeros((tun.n, 4), dtype=torch.float32, device='cuda')
    with pytest.raises(ValueError):
        tun.step_advect(0.05, 0.52, 0.16, 0, other[:-1], pos, spd, ids, 1.0, 100.0, 1, BOX)
    with pytest.raises(ValueError):
        tun.set_mac(other.double())
    tun.particle_gather(pos, spd, ids, order, *out)
    tun.pack_particles_slots(pos, spd, ids, q16, q8)
    torch.cuda.synchronize()
    assert torch.equal(out[0], pos)


def test_slice_and_cp_buffers_are_checked(tun):
    nx, ny, nz = DIMS
    with pytest.raises(ValueError):
        tun.slice(torch.zeros(2 * nx * ny - 1, dtype=torch.uint8, device='cuda'), 2, 5, 0, 0.05, 4.0, 0.0, 1.5)
    with pytest.raises(ValueError):
        tun.slice(torch.zeros(2 * nx * nz, dtype=torch.uint8, device='cuda'), 3, 5, 0, 0.05, 4.0, 0.0, 1.5)
    assert tun.slice_dims(0) == (nz, ny) and tun.slice_dims(1) == (nx, nz)
    big = torch.zeros(2 * nx * nz, dtype=torch.uint8, device='cuda')
    tun.slice(big, 1, 5, 0, 0.05, 4.0, 0.0, 1.5)             # larger buffers are fine (lbm.py's slice_buf)
    verts = torch.full((6, 3), 5.0, dtype=torch.float32, device='cuda')
    norms = torch.zeros((6, 3), dtype=torch.float32, device='cuda')
    with pytest.raises(ValueError):
        tun.cp(verts, norms[:5], torch.zeros(6, dtype=torch.int8, device='cuda'), 0.05, 1.2)
    with pytest.raises(ValueError):
        tun.cp(verts, norms, torch.zeros(6, dtype=torch.uint8, device='cuda'), 0.05, 1.2)
    tun.cp(verts, norms, torch.zeros(6, dtype=torch.int8, device='cuda'), 0.05, 1.2)
    torch.cuda.synchronize()
