import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _has_gpu():
        pytest.skip("no GPU")
    import torch
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def geometry():
    from cuda_hlt_amd import gen
    return gen.geometry()


@pytest.fixture(scope="session")
def module_zs():
    from cuda_hlt_amd import gen
    return gen.module_zs()


@pytest.fixture(scope="session")
def minbias_1000():
    """1000 nu=7.6 minimum-bias events, seed 0 (the CPU parity sample)."""
    from cuda_hlt_amd import gen
    return gen.minbias(1000, seed=0)


@pytest.fixture(scope="session")
def minbias_200():
    from cuda_hlt_amd import gen
    return gen.minbias(200, seed=0)
