"""The workload session (windtunnel.session.Session, the bench's config-2 loop):
re-sorting the particle slots by cell changes nothing but speed -- every
particle, the populations, forces and slices equal those of a session whose
slots are sorted only once -- and fusing each step's particle advection into
the next step's launch changes nothing either."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(sort_every, dims, mask, steps, fuse=True, form='fp32', peek=()):
    from windtunnel.session import Session
    s = Session(dims, mask, form, 0.05, 0.52, 0.16, particles=20011, slices=[(2, dims[2] // 2, 0), (1, 9, 2)],
                frame_every=5, lref=12.0, seed=3, fuse=fuse)
    s.sort_every = sort_every
    for k in range(steps):
        s.step()
        if k in peek:                                     # read the particles mid-run (flushes)
            s.particles_by_id()
    torch.cuda.synchronize()
    pos, spd = s.particles_by_id()
    return dict(f=s.T.raw_populations().cpu().numpy(), mac=s.T.mac.cpu().numpy(), sums=s.T.sums.cpu().numpy(),
                forces=s.forces_host.numpy().copy(), slices=[h.numpy().copy() for h in s.slice_host],
                pos=pos, spd=spd, ids=s.ids.cpu().numpy())


def _mask(dims):
    from windtunnel import geometry
    ells = geometry.car_body(dims, 5, count=5, box=((25, 45), (1, 12), (12, 24)), axes=(2.0, 6.0))
    return geometry.ellipsoids(ells, dims, ground=True)[0]


def _same(a, b):
    for k in ('f', 'mac', 'sums', 'forces', 'pos', 'spd', 'ids'):
        assert a[k].tobytes() == b[k].tobytes(), k
    for x, y in zip(a['slices'], b['slices']):
        assert np.array_equal(x, y)


def test_resorted_session_equals_sorted_once(gpu):
    dims = (96, 40, 36)
    mask = _mask(dims)
    a = _run(7, dims, mask, 23)
    b = _run(10 ** 9, dims, mask, 23)
    assert not np.array_equal(a['ids'], b['ids'])        # the slots really were re-sorted
    for k in ('f', 'mac', 'sums', 'forces', 'pos', 'spd'):
        assert a[k].tobytes() == b[k].tobytes(), k
    for x, y in zip(a['slices'], b['slices']):
        assert np.array_equal(x, y)
    assert np.isfinite(a['pos']).all() and (a['spd'] > 0).all()


@pytest.mark.parametrize('form', ['fp32', 'fp16', 'bf16', 'fp64'])
def test_fused_advection_equals_separate_launches(gpu, form):
    """wt_step_advect (step k + 1 fused with the advection of step k, two
    alternating mac buffers) gives what wt_step followed by wt_advect_slots
    gives, bit for bit, including mid-run reads that flush the pending step."""
    dims = (96, 40, 36)
    mask = _mask(dims)
    a = _run(7, dims, mask, 23, fuse=True, form=form, peek=(4, 5, 11))
    b = _run(7, dims, mask, 23, fuse=False, form=form)
    _same(a, b)
    assert np.isfinite(a['pos']).all() and (a['spd'] > 0).all()


@pytest.mark.parametrize('form', ['fp32', 'fp16'])
def test_fused_advection_with_more_particle_blocks_than_tiles(gpu, form):
    """A 24x16x12 tunnel (18 tiles of 256 cells) with 20011 particles (79
    workgroups of them): the fused launch's grid holds more advection workgroups
    than tiles, so nothing is spread between them (its first 79 workgroups
    advect and the 18 tiles follow). Results equal the separate launches bit
    for bit."""
    from windtunnel import geometry
    dims = (24, 16, 12)
    mask = geometry.sphere(dims, (8, 8, 6), 3)[0]
    a = _run(7, dims, mask, 13, fuse=True, form=form, peek=(4,))
    b = _run(7, dims, mask, 13, fuse=False, form=form)
    _same(a, b)
    assert np.isfinite(a['pos']).all()


def test_step_advect_rejects_aliased_mac(gpu):
    """The advected field must not be the buffer the same launch writes."""
    from windtunnel.session import Session
    dims = (40, 20, 20)
    s = Session(dims, np.zeros(dims[::-1], np.uint8), 'fp32', particles=64, fuse=True)
    with pytest.raises(RuntimeError, match='mac'):
        s.T.step_advect(0.05, 0.52, 0.16, 0, s.T.mac, s.pos, s.spd, s.ids, 1.0, 100.0, 1, s.box)


def test_roofline_times_fused_launches(gpu):
    """windtunnel.roofline's fused timing path (the kernel the counter runs price)."""
    from windtunnel import roofline
    dims = (64, 32, 32)
    t, r = roofline.step_timing('fp32', dims, np.zeros(dims[::-1], np.uint8), steps=3, warmup=2, particles=1000)
    assert r['what'] == 'step_advect' and r['bytes_per_launch'] == 169 * t.n + 40 * 1000
    assert r['ms_median'] > 0 and t.mac.data_ptr() == t.grid.mac


def test_viewer_average_equals_frame_by_frame_reads(gpu):
    """Session.viewer_coefficients (forces logged on the device each frame) gives
    exactly the viewer's average of the forces read synchronously frame by frame."""
    from windtunnel.coefficients import ForceAverage
    from windtunnel.session import FRAME_LOG, Session
    from windtunnel import geometry
    dims = (48, 20, 20)
    mask = geometry.sphere(dims, (14, 10, 10), 6)[0]
    s = Session(dims, mask, 'fp32', 0.05, 0.52, 0.16, frame_every=5, frontal=113)
    ref = ForceAverage(0.05, 113)
    for _ in range(1000):
        before = s.frames
        s.step()
        if s.frames != before:
            torch.cuda.synchronize()
            f = s.forces_host.numpy()
            ref.add(f[0], f[1], 5)
    assert s.frames == 200 and ref.settled
    assert s.viewer_coefficients() == (ref.cd, ref.cl)
    t = Session(dims, mask, 'fp32', 0.05, 0.52, 0.16, frame_every=1, frontal=113)
    for _ in range(FRAME_LOG + 1):
        t.step()
    with pytest.raises(RuntimeError, match='log holds'):
        t.viewer_coefficients()
