"""wt_step_noalloc -- the step whose force/mass reduction slots live in the
library's static device scratch instead of the grid's partials buffer -- equals
wt_step bit for bit (populations, macroscopic field and FP64 sums), on two calls
in a row, the second on a smaller tunnel, through the C ABI."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize('form', ['fp32', 'fp16'])
def test_noalloc_step_equals_wt_step(gpu, form):
    from windtunnel import _lib, geometry
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS, Params
    from windtunnel.tunnel import Tunnel
    lib = _lib.load()
    flags = WT_FLAG_FORCES | WT_FLAG_MASS
    for dims, steps in (((160, 64, 72), 7), ((48, 20, 22), 4)):      # the second call is smaller
        mask, _ = geometry.sphere(dims, (dims[0] // 3, dims[1] // 2, dims[2] // 2), dims[1] // 3)
        ref, nal = Tunnel(dims, form), Tunnel(dims, form)
        for t in (ref, nal):
            t.init(0.05)
            t.set_solid(mask)
        nal.grid.partials = None                 # no reduction buffer: the scratch serves
        ref.step(steps, 0.05, 0.52, 0.16, flags=flags)
        p = Params(0.05, 0.52, 0.16)
        rc = lib.wt_step_noalloc(nal.gp, nal.cur, steps, ctypes.byref(p), flags, _lib.stream_handle())
        assert rc == 0, _lib.last_error()
        nal.cur ^= steps & 1
        torch.cuda.synchronize()
        assert _same(ref.raw_populations().cpu().numpy(), nal.raw_populations().cpu().numpy()), (form, dims)
        assert _same(ref.mac.cpu().numpy(), nal.mac.cpu().numpy()), (form, dims)
        assert _same(ref.sums.cpu().numpy(), nal.sums.cpu().numpy()), (form, dims)
