import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def wg():
    import watergap2_amd
    return watergap2_amd


@pytest.fixture(scope="session")
def oracle():
    from oracle import wg_oracle
    return wg_oracle


@pytest.fixture(scope="session")
def regional_world(wg):
    """The 2,000-cell 50 x 40 regional grid, seed 0."""
    return wg.World(wg.regional_spec(seed=0))


@pytest.fixture(scope="session")
def regional_edge_world(wg):
    """Regional grid with cells that take rarely used branches."""
    return wg.World(wg.regional_spec(seed=5, nrows=30, ncols=40, edge_cases=True, glolake_every=20,
                                     res_every=25))


@pytest.fixture(scope="session")
def global_world(wg):
    """The 67,420-cell global grid, seed 1."""
    return wg.World(wg.global_spec(seed=1))


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
