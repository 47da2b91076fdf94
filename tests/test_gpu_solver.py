"""The drop-in lbm.Solver driven the way server/app.py drives it, checked
end to end against the oracle replaying the same sequence (lbm.py:398-482)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P = dict(uIn=0.1, tau=0.51, smag=0.16, steps=10, maxAge=1000.0, lref=40.0,
         seedX=2.0, seedCy=20.0, seedCz=20.0, seedHy=10.0, seedHz=10.0,
         slice=None, cp=False, running=True, particles=5000)   # app.py:62-64


def _oracle_particles(solver, count, p, seed):
    rng = np.random.default_rng(seed)
    d = np.zeros((count, 4), np.float32)
    d[:, 0] = 2 + rng.random(count) * (solver.nx - 8)
    d[:, 1] = p['seedCy'] + (rng.random(count) * 2 - 1) * p['seedHy']
    d[:, 2] = p['seedCz'] + (rng.random(count) * 2 - 1) * p['seedHz']
    d[:, 3] = rng.random(count) * p['maxAge'] * 0.5
    return d, np.ones(count, np.float32)


def test_solver_frame_matches_oracle(gpu):
    import lbm
    from windtunnel import geometry
    dims = lbm.RESOLUTIONS['test']               # (64, 28, 36)
    s = lbm.Solver(dims, lbm.get_preferred_device(), seed=123)
    s.reset(P['uIn'])
    mask, counts = geometry.sphere(dims, (20, 13, 17), 10, ground=True)
    s.set_solid(mask)
    s.reset(P['uIn'])
    rng = np.random.default_rng(1)
    verts = np.column_stack([rng.uniform(14, 26, 500), rng.uniform(8, 18, 500), rng.uniform(12, 22, 500)])
    norms = rng.standard_normal((500, 3))
    norms /= np.linalg.norm(norms, axis=1, keepdims=True)
    s.set_surface(verts.astype(np.float32), norms.astype(np.float32))
    p = dict(P, seedCy=13.0, seedCz=17.0, seedHy=8.0, seedHz=8.0, maxAge=300.0)
    s.set_particles(2000, p)
    s.sort_every = 2                              # re-sort the particle slots inside the test

    o = O.OracleTunnel(dims, O.FP32)
    o.init(p['uIn'])
    o.set_solid(mask)
    o.init(p['uIn'])
    pos, spd = _oracle_particles(s, 2000, p, 123)
    seed = 1
    for frame in range(6):
        s.step(p['steps'], p)
        o.step(p['steps'], p['uIn'], p['tau'], p['smag'])
        seed += 1
        pos, spd = O.advect(pos, spd, o.mac, mask, dims, float(p['steps']), p['uIn'], p['maxAge'], seed,
                            (p['seedX'], p['seedCy'], p['seedCz'], p['seedHy'], p['seedHz']))
        pf = dict(p, slice={'axis': 2, 'index': 17, 'field': frame % 4, 'vmin': -1.2, 'vmax': 1.6})
        out = s.frame(pf, want_cp=True)
        assert set(out) == {'parts16', 'parts8', 'slice', 'sliceW', 'sliceH', 'sliceAxis', 'cp', 'forces'}
        assert out['parts16'].dtype == np.uint16 and out['parts16'].shape == (4 * 2000,)
        assert out['parts8'].dtype == np.uint8 and out['parts8'].shape == (2000,)
        assert (out['sliceW'], out['sliceH'], out['sliceAxis']) == (64, 28, 2)
        assert out['slice'].shape == (2 * 64 * 28,) and out['cp'].dtype == np.int8
        assert isinstance(out['forces'], list) and len(out['forces']) == 3
        q16, q8 = O.pack_particles(pos, spd, dims)
        assert np.array_equal(out['parts16'], q16) and np.array_equal(out['parts8'], q8)
        sl = O.slice_(o.mac, mask, dims, 2, 17, frame % 4, p['uIn'], p['lref'], -1.2, 1.6)
        assert np.array_equal(out['slice'], sl)
        cpv = O.cp(o.mac, mask, verts, norms, dims, p['uIn'], lbm.CP_RANGE)
        assert np.array_equal(out['cp'], cpv)
        np.testing.assert_allclose(out['forces'], o.forces(), rtol=1e-12, atol=1e-15)
    # Cd from the frontal area as main.js:336-339 does
    q = 0.5 * p['uIn'] ** 2 * max(1, counts['frontal'])
    cd = out['forces'][0] / q
    assert np.isfinite(cd) and abs(cd) < 1e3


def test_forces_after_reset_without_steps(gpu):
    import lbm
    from windtunnel import geometry
    dims = (40, 20, 20)
    s = lbm.Solver(dims)
    mask, _ = geometry.sphere(dims, (15, 10, 10), 6)
    s.set_solid(mask)
    s.reset(0.05)
    out = s.frame(dict(P, uIn=0.05), want_cp=False)
    o = O.OracleTunnel(dims, O.FP32)
    o.set_solid(mask)
    o.init(0.05)
    np.testing.assert_allclose(out['forces'], o.forces(), rtol=1e-12, atol=1e-15)
    assert 'parts16' not in out and 'slice' not in out and 'cp' not in out


def test_module_surface(gpu):
    import lbm
    assert lbm.BYTES_PER_CELL == 169 and lbm.CP_RANGE == 1.2
    assert lbm.RESOLUTIONS['max'] == (544, 224, 272)
    s = lbm.Solver(lbm.RESOLUTIONS['test'], None)
    assert (s.nx, s.ny, s.nz, s.n, s.num_p, s.num_verts) == (64, 28, 36, 64 * 28 * 36, 0, 0)
    assert s.slice_dims(0) == (36, 28) and s.slice_dims(1) == (64, 36) and s.slice_dims(2) == (64, 28)
    with pytest.raises(AssertionError):
        s.set_solid(np.zeros(5, np.uint8))
    dev = lbm.get_preferred_device()
    assert dev.is_cuda and dev.total_memory > 0
