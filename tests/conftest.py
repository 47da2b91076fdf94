import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, 'wind-tunnel_amd')
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP GPU (MI355X); runs the libwt kernels')
    config.addinivalue_line('markers', 'slow: long-running check')


@pytest.fixture(scope='session')
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope='session')
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail('GPU test selected but torch.cuda.is_available() is False')
    import windtunnel
    windtunnel.load(build_if_missing=False)
    return torch.device('cuda:0')
