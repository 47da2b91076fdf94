"""Continuous batching: sessions admitted while others run, memory-capped
admission, and every session's per-step force stream bit-identical alone and
in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests():
    from windtunnel.serving import mixed_workload
    return mixed_workload(seed=11, duration=1.5, rate=8.0, min_cells=40e3, max_cells=600e3, steps=(20, 60),
                          max_sessions=8)


def _alone(req, slice_step=0):
    """Per-step forces of the session run alone (every step storing its field),
    and the serving loop's slice image after step slice_step."""
    import torch
    from windtunnel import geometry
    from windtunnel._lib import WT_FLAG_FORCES
    from windtunnel.session import FIELD_RANGES
    from windtunnel.tunnel import Tunnel
    mask, _ = geometry.ellipsoids(req.ellipsoids, req.dims, ground=req.ground)
    t = Tunnel(req.dims, req.form)
    t.init(req.uin)
    t.set_solid(mask)
    out, img = [], None
    nx, ny, nz = req.dims
    for k in range(req.steps):
        t.step(1, req.uin, req.tau, req.smag, flags=WT_FLAG_FORCES)
        out.append(t.sums.cpu().numpy()[:3].copy())
        if k + 1 == slice_step:
            dev = torch.zeros(2 * nx * ny, dtype=torch.uint8, device='cuda')
            t.slice(dev, 2, nz // 2, 0, req.uin, 40.0, *FIELD_RANGES[0])
            img = dev.cpu().numpy()
    return (np.array(out), img) if slice_step else np.array(out)


@pytest.mark.parametrize('graphs', [False, True])
def test_sessions_identical_alone_and_in_the_mix(gpu, graphs):
    from windtunnel.serving import Server
    from windtunnel.tunnel import state_bytes
    reqs = _requests()
    assert len(reqs) >= 6 and len({r.form for r in reqs}) == 2
    biggest = max(state_bytes(r.dims, r.form) + 4 * int(np.prod(r.dims)) for r in reqs)
    srv = Server(budget_bytes=int(2.5 * biggest), use_graphs=graphs)   # forces queueing
    st = srv.run(reqs)
    assert st['sessions_done'] == len(reqs)
    assert st['peak_resident_gb'] * 1e9 <= 2.5 * biggest
    assert max(srv.queue_delays) > 0.0          # someone waited for memory
    assert st['step_latency_ms']['p50'] <= st['step_latency_ms']['p99']
    assert st['cell_steps'] == sum(r.steps * int(np.prod(r.dims)) for r in reqs)
    for r in reqs:
        lv = srv.finished[r.sid]
        assert lv.done_steps == r.steps and len(lv.forces) == r.steps
        got = np.array(lv.forces)
        assert lv.slices == r.steps // r.slice_every and lv.slices > 0
        # the field is stored only on the steps whose slice is due (WT_FLAG_NO_MAC with
        # gates): the last slice equals the one taken alone after the same step
        want, img = _alone(r, lv.last_slice_step)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), r.sid
        assert np.array_equal(lv.slice_host.numpy(), img), (r.sid, lv.last_slice_step)


def test_preemption_keeps_results_exact(gpu):
    """Three equal sessions arrive together and only two fit: the third evicts the
    one with the most remaining work to host memory; the evicted session later
    resumes from its checkpoint. Every force stream equals the session run alone."""
    from windtunnel.serving import Request, Server
    from windtunnel.tunnel import state_bytes
    dims = (96, 48, 48)
    reqs = [Request(k, 0.001 * k, dims, np.array([[30.0, 10.0 + 3 * k, 24.0, 8.0, 6.0, 6.0]]),
                    'fp32' if k != 1 else 'fp16', 120 + 40 * k, 0.05 + 0.01 * k, 0.52, 0.16, True, 20)
            for k in range(3)]
    one = max(state_bytes(dims, r.form) + 4 * int(np.prod(dims)) for r in reqs)
    srv = Server(budget_bytes=int(2.5 * one), preempt_after=0.0)
    st = srv.run(reqs)
    assert st['sessions_done'] == 3
    assert st['evictions'] >= 1 and any(srv.finished[k].evictions for k in range(3))
    for r in reqs:
        lv = srv.finished[r.sid]
        assert lv.done_steps == r.steps and len(lv.forces) == r.steps
        assert np.array_equal(np.array(lv.forces).view(np.uint64), _alone(r).view(np.uint64)), r.sid


def test_tick_target_bounds_the_active_set(gpu):
    """With a tick-time target (fixed rate estimate, so the test is deterministic)
    the active set never exceeds the target's cost units unless it is a single
    session, everything completes, and every force stream is still exact."""
    from windtunnel.serving import Server, TickBudget
    reqs = _requests()
    units = [TickBudget.units(r.dims, r.form) for r in reqs]
    # room for about two average sessions per tick
    tb = TickBudget(1.0, sec_per_unit=1e-3 / (2.0 * float(np.mean(units))), adapt=False)
    srv = Server(tick_budget=tb)
    st = srv.run(reqs)
    assert st['sessions_done'] == len(reqs)
    limit = tb.target / tb.spu
    assert st['max_tick_units'] <= max(limit, max(units)) + 1e-6
    assert st['max_tick_units'] < sum(units)            # the target really held sessions back
    assert max(srv.queue_delays) > 0.0
    for r in reqs:
        lv = srv.finished[r.sid]
        assert np.array_equal(np.array(lv.forces).view(np.uint64), _alone(r).view(np.uint64)), r.sid
