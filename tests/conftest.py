import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-size workloads")


def _ensure_built():
    from mkf_amd import LIB_PATH
    from mkf_amd import build
    if not os.path.exists(LIB_PATH):
        build.build_library()
    if not os.path.exists(build.ORACLE_LIB):
        build.build_oracle()


_ensure_built()


@pytest.fixture(scope="session")
def lib():
    import mkf_amd
    return mkf_amd.load()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    oracle.set_threads(min(16, os.cpu_count() or 1))
    return oracle


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no GPU is visible")
    return torch.device("cuda:0")
