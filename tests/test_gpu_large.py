"""Size limits and empty inputs through the C ABI.

* A 640x320x288 FP32 tunnel (59M cells: 19*n*4 B of populations exceed 32-bit
  buffer offsets, so the step kernel takes its 64-bit global-address form on
  its own) matches the oracle bit for bit over two steps, forces within rtol
  1e-12 (different summation order).
* Zero particles, vertices or steps are no-ops that return success."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_large_tunnel_global_addressing_bit_exact(gpu):
    from windtunnel import geometry
    from windtunnel._lib import WT_FLAG_FORCES
    from windtunnel.tunnel import Tunnel
    dims = (640, 320, 288)
    n = dims[0] * dims[1] * dims[2]
    assert 19 * n * 4 >= 1 << 32          # beyond 32-bit byte offsets
    ells = geometry.car_body(dims, 3, count=6, box=((150, 300), (1, 90), (110, 180)), axes=(8.0, 30.0))
    mask = geometry.ellipsoids(ells, dims, ground=True)[0]
    t = Tunnel(dims, 'fp32')
    t.init(0.05)
    t.set_solid(mask)
    t.step(2, 0.05, 0.51, 0.16, flags=WT_FLAG_FORCES)
    gpu_f = t.raw_populations().cpu().numpy()
    gpu_mac = t.mac.cpu().numpy()
    gpu_sums = t.sums.cpu().numpy()
    del t
    torch.cuda.empty_cache()
    O.set_threads(16)
    o = O.OracleTunnel(dims, O.FP32)
    o.init(0.05)
    o.set_solid(mask)
    o.step(2, 0.05, 0.51, 0.16)
    assert np.array_equal(gpu_f.view(np.uint32), o.raw().view(np.uint32))
    assert np.array_equal(gpu_mac.reshape(-1).view(np.uint32), o.mac.reshape(-1).view(np.uint32))
    ref = o.forces()
    assert np.allclose(gpu_sums[:3], ref[:3], rtol=1e-12, atol=0)


def test_empty_inputs_are_noops(gpu):
    from windtunnel.tunnel import Tunnel
    dims = (32, 12, 12)
    t = Tunnel(dims, 'fp32')
    t.init(0.05)
    t.set_solid(np.zeros(dims[0] * dims[1] * dims[2], np.uint8))
    before = t.raw_populations().cpu().numpy().copy()
    t.step(0, 0.05, 0.52, 0.16)
    assert np.array_equal(before, t.raw_populations().cpu().numpy())
    pos = torch.empty((0, 4), dtype=torch.float32, device='cuda')
    spd = torch.empty(0, dtype=torch.float32, device='cuda')
    ids = torch.empty(0, dtype=torch.int32, device='cuda')
    box = (2.0, 6.0, 6.0, 3.0, 3.0)
    t.advect(pos, spd, 1.0, 0.05, 10.0, 1, box)
    t.advect_slots(pos, spd, ids, 1.0, 0.05, 10.0, 1, box)
    q16 = torch.empty(0, dtype=torch.int16, device='cuda')
    q8 = torch.empty(0, dtype=torch.uint8, device='cuda')
    t.pack_particles(pos, spd, q16, q8)
    verts = torch.empty((0, 3), dtype=torch.float32, device='cuda')
    cp = torch.empty(0, dtype=torch.int8, device='cuda')
    t.cp(verts, verts, cp, 0.05, 1.2)
    torch.cuda.synchronize()
