import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "cuda-ipc-benchmarks_amd"))
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def have_gpu():
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def amd():
    import lulesh_amd
    lulesh_amd.init()
    return lulesh_amd


def bits(a):
    import numpy as np
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def assert_bitwise(a, b, what=""):
    import numpy as np
    ba, bb = bits(a), bits(b)
    if not np.array_equal(ba, bb):
        diff = np.nonzero(ba != bb)[0]
        i = diff[0]
        raise AssertionError(f"{what}: {diff.size} of {ba.size} values differ; first at {i}: "
                             f"{a[i]!r} vs {b[i]!r}")
