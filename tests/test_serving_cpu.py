"""Host-side checks of the serving workload generator (no GPU)."""
import numpy as np


def test_mixed_workload_shape():
    from windtunnel.serving import mixed_workload
    reqs = mixed_workload(seed=3, duration=600.0, rate=2.0)
    n = len(reqs)
    assert 1000 < n < 1400                       # Poisson(1200)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0) and arr[-1] <= 600.0
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.5) < 0.05         # mean inter-arrival 1/rate
    cells = np.array([np.prod(r.dims) for r in reqs], dtype=float)
    for r in reqs:
        nx, ny, nz = r.dims
        assert nx == 2 * ny and ny == nz and ny % 8 == 0
        assert 200 <= r.steps <= 5000
        assert 0.03 <= r.uin <= 0.08 and 0.51 <= r.tau <= 0.60
    assert cells.min() >= 0.3e6 and cells.max() <= 40e6
    # log-uniform: median near the geometric mean of the range
    assert 2e6 < np.median(cells) < 8e6
    assert [r.form for r in reqs[:4]] == ['fp32', 'fp16', 'fp32', 'fp16']
    again = mixed_workload(seed=3, duration=600.0, rate=2.0)
    assert [r.dims for r in again] == [r.dims for r in reqs]


def test_tick_budget_admission_rule():
    from windtunnel.serving import FORM_COST, TickBudget
    tb = TickBudget(10.0, sec_per_unit=1e-9, adapt=False)        # 10 ms = 1e7 units
    assert tb.units((100, 100, 100), 'fp32') == 1e6
    assert tb.units((100, 100, 100), 'fp16') == 1e6 * 85 / 169
    assert FORM_COST['fp64'] == 321 / 169
    assert tb.admits(0.0, 5e8)                                    # an idle GPU takes any head
    assert tb.admits(9e6, 1e6) and not tb.admits(9e6, 1.5e6)
    tb.observe(0.02, 1e7)                                         # adapt=False: unchanged
    assert tb.spu == 1e-9


def test_tick_budget_learns_median_rate():
    from windtunnel.serving import TickBudget
    tb = TickBudget(50.0, sec_per_unit=1.0, window=5)
    for dt in (0.010, 0.011, 0.5, 0.009, 0.010):                  # one outlier (a rebuild)
        tb.observe(dt, 1e7)
    assert abs(tb.spu - 1.0e-9) < 1e-15                           # median, not mean
    tb.observe(0.0, 1e7)                                          # empty samples are ignored
    tb.observe(0.01, 0.0)
    assert len(tb.samples) == 5
    assert tb.predicted(2e7) == 2e7 * tb.spu


def test_prefilled_workload():
    """prefill=N adds N sessions resident at t = 0 (steps still to run drawn over
    a length-biased session length) in front of the unchanged-in-kind Poisson
    stream; the plain stream (no prefill) is the recorded one."""
    import numpy as np
    from windtunnel.serving import mixed_workload
    plain = mixed_workload(seed=0, duration=600.0, rate=2.0)
    assert len(plain) == 1183 and plain[0].dims == (176, 88, 88) and plain[0].steps == 396
    reqs = mixed_workload(seed=0, duration=30.0, rate=2.0, prefill=40)
    assert all(r.arrival == 0.0 for r in reqs[:40]) and all(r.arrival > 0.0 for r in reqs[40:])
    assert all(1 <= r.steps <= 5000 for r in reqs[:40]) and all(200 <= r.steps <= 5000 for r in reqs[40:])
    assert [r.sid for r in reqs] == list(range(len(reqs)))
    cells = np.array([np.prod(r.dims) for r in reqs])
    assert cells.min() >= 0.4e6 and cells.max() <= 34e6
    assert len({r.form for r in reqs[:40]}) == 2
