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
