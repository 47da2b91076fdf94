"""Storage precision against the FP64 form on a developed flow: the 128x64x64
sphere (D = 16, uin 0.05, tau 0.52, Cs 0.17; Re 120, a steady wake) with the
force every step, Cd and Cl averaged over steps 6000-8000 (after 2.3
flow-throughs of 2,560 steps).

Stated tolerances on the time-averaged Cd relative to FP64's: FP32 0.1 %,
FP16-deviation 1 % (the stated target for the reduced-precision form;
measured 0.17 %), BF16-deviation 10 % -- BF16's 8-bit significand misses the
1 % bar (measured 6 %), so its bound only guards against a regression. |dCl|
(Cl ~ 0 by symmetry): FP32 1e-5, FP16 1e-3."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_time_averaged_cd_cl_against_fp64(gpu):
    from windtunnel import precision
    r = precision.coefficients(scale=1, avg_from=6000, avg_to=8000)['forms']
    assert all(v['finite'] for v in r.values()), r
    assert 1.1 < r['fp64']['cd'] < 1.4, r['fp64']          # a sphere's drag at Re ~ 100
    assert r['fp32']['cd_rel_to_fp64'] < 1e-3, r['fp32']
    assert r['fp16']['cd_rel_to_fp64'] < 1e-2, r['fp16']
    assert r['bf16']['cd_rel_to_fp64'] < 1e-1, r['bf16']
    assert abs(r['fp32']['cl_minus_fp64']) < 1e-5, r['fp32']
    assert abs(r['fp16']['cl_minus_fp64']) < 1e-3, r['fp16']
