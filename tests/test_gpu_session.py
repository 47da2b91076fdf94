"""The workload session (windtunnel.session.Session, the bench's config-2 loop):
re-sorting the particle slots by cell changes nothing but speed -- every
particle, the populations, forces and slices equal those of a session whose
slots are sorted only once."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(sort_every, dims, mask, steps):
    from windtunnel.session import Session
    s = Session(dims, mask, 'fp32', 0.05, 0.52, 0.16, particles=20011, slices=[(2, dims[2] // 2, 0), (1, 9, 2)],
                frame_every=5, lref=12.0, seed=3)
    s.sort_every = sort_every
    for _ in range(steps):
        s.step()
    torch.cuda.synchronize()
    pos, spd = s.particles_by_id()
    return dict(f=s.T.raw_populations().cpu().numpy(), mac=s.T.mac.cpu().numpy(), sums=s.T.sums.cpu().numpy(),
                forces=s.forces_host.numpy().copy(), slices=[h.numpy().copy() for h in s.slice_host],
                pos=pos, spd=spd, ids=s.ids.cpu().numpy())


def test_resorted_session_equals_sorted_once(gpu):
    from windtunnel import geometry
    dims = (96, 40, 36)
    ells = geometry.car_body(dims, 5, count=5, box=((25, 45), (1, 12), (12, 24)), axes=(2.0, 6.0))
    mask = geometry.ellipsoids(ells, dims, ground=True)[0]
    a = _run(7, dims, mask, 23)
    b = _run(10 ** 9, dims, mask, 23)
    assert not np.array_equal(a['ids'], b['ids'])        # the slots really were re-sorted
    for k in ('f', 'mac', 'sums', 'forces', 'pos', 'spd'):
        assert a[k].tobytes() == b[k].tobytes(), k
    for x, y in zip(a['slices'], b['slices']):
        assert np.array_equal(x, y)
    assert np.isfinite(a['pos']).all() and (a['spd'] > 0).all()
