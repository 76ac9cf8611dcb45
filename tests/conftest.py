import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPU tests need a HIP device (torch.cuda.is_available() is False)")
    import lig_amd._lib as L
    L.load()
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _bounds_checks():
    """With LIG_CHECK_ASSERT=1 (and LIG_LIB_VARIANT=checked), every test that
    used the library must leave the bounds-checked build's failure count at
    zero."""
    yield
    if os.environ.get("LIG_CHECK_ASSERT") != "1":
        return
    import lig_amd._lib as L
    if L._lib is None:   # the test did not load the library
        return
    st = L.check_status()
    assert st is not None, f"{L.LIB_PATH} is not the bounds-checked build"
    failures, line, path = st
    assert failures == 0, f"{failures} failed bounds checks, first at {path}:{line}"
