"""Session checkpoint / resume: evicting a tunnel to host memory and restoring
it continues bit for bit as if it had never left the GPU."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('form', ['fp32', 'fp16'])
def test_checkpoint_resume_is_exact(gpu, form):
    from windtunnel import geometry
    from windtunnel._lib import WT_FLAG_FORCES
    from windtunnel.tunnel import Tunnel
    dims = (56, 24, 22)
    mask, _ = geometry.ellipsoids([[18, 8, 11, 6, 5, 5]], dims, ground=True)
    a = Tunnel(dims, form)
    a.init(0.06)
    a.set_solid(mask)
    a.step(7, 0.06, 0.53, 0.16)                 # odd: cur = 1
    st = a.checkpoint()
    torch.cuda.synchronize()
    a.step(9, 0.06, 0.53, 0.16, flags=WT_FLAG_FORCES)
    b = Tunnel.from_checkpoint(st)
    b.step(9, 0.06, 0.53, 0.16, flags=WT_FLAG_FORCES)
    torch.cuda.synchronize()
    fa, fb = a.raw_populations().cpu().numpy(), b.raw_populations().cpu().numpy()
    assert np.array_equal(fa.view(np.uint8), fb.view(np.uint8))
    assert np.array_equal(a.mac.cpu().numpy().view(np.uint32), b.mac.cpu().numpy().view(np.uint32))
    assert np.array_equal(a.sums.cpu().numpy(), b.sums.cpu().numpy())
