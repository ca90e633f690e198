import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_generation.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as d:
        return {k: d[k] for k in d.keys()}


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def params_from(arr):
    from nanoxgen_amd import abi
    return abi.GenerationParams.from_buffer_copy(arr.tobytes())


def aligned_copy(blob):
    from nanoxgen_amd import abi
    out = abi.aligned_bytes(blob.nbytes)
    out[:] = blob
    return out
