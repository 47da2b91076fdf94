"""The viewer's Cd/Cl average (windtunnel/coefficients.py, web/js/main.js:336-351),
checked against what its definition fixes: settling, the running mean while the
window fills, the exponential window after, the area floor and divergence."""
import math

import numpy as np

from windtunnel.coefficients import SETTLE_STEPS, WINDOW_STEPS, ForceAverage, dynamic_pressure


def test_dynamic_pressure_floors_the_area_at_one():
    assert dynamic_pressure(0.05, 0) == 0.5 * 0.05 * 0.05
    assert dynamic_pressure(0.05, 197) == 0.5 * 0.05 * 0.05 * 197


def test_settling_then_running_mean_then_exponential_window():
    uin, area, dt = 0.05, 200, 10
    q = dynamic_pressure(uin, area)
    avg = ForceAverage(uin, area)
    rng = np.random.default_rng(1)
    fx = rng.normal(1.0, 0.2, 600)
    fy = rng.normal(0.0, 0.1, 600)
    out = [avg.add(a, b, dt) for a, b in zip(fx, fy)]
    first = SETTLE_STEPS // dt - 1                  # frame index at which 800 steps are reached
    assert all(o is None for o in out[:first]) and out[first] is not None
    # while w < 2500 the average is the arithmetic mean of the settled frames
    k_full = first + WINDOW_STEPS // dt - 1
    for k in (first, first + 7, k_full):
        mean = fx[first:k + 1].mean() / q
        assert math.isclose(out[k][0], mean, rel_tol=1e-12)
    # after that each frame moves it by dt / 2500: an exponential window
    cd = out[k_full][0]
    a = dt / WINDOW_STEPS
    for k in range(k_full + 1, 600):
        cd = cd + (fx[k] / q - cd) * a
    assert math.isclose(out[-1][0], cd, rel_tol=1e-12)
    assert avg.w == WINDOW_STEPS and avg.steps == 600 * dt


def test_constant_force_gives_its_coefficient_and_uneven_frames_weigh_by_steps():
    avg = ForceAverage(0.08, 50)
    q = dynamic_pressure(0.08, 50)
    for steps in (300, 500, 7, 1000, 2000, 3):
        r = avg.add(1.5 * q, -0.25 * q, steps)
    assert math.isclose(r[0], 1.5, rel_tol=1e-14) and math.isclose(r[1], -0.25, rel_tol=1e-14)
    b = ForceAverage(0.05, 1)
    b.add(0.0, 0.0, 800)                              # settles on its first frame: cd 0 with weight 800
    r = b.add(dynamic_pressure(0.05, 1), 0.0, 200)    # cd 1 with weight 200
    assert math.isclose(r[0], 200 / 1000, rel_tol=1e-14)


def test_divergence_stops_the_average():
    avg = ForceAverage(0.05, 100)
    q = dynamic_pressure(0.05, 100)
    avg.add(q, 0.0, 900)
    assert avg.add(float('nan'), 0.0, 10) is None and avg.diverged
    assert avg.add(q, 0.0, 10) is None
    big = ForceAverage(0.05, 100)
    assert big.add(1001 * q, 0.0, 900) is None and big.diverged
    assert ForceAverage(0.05, 100).add(999 * q, 0.0, 900) is not None
