import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "photonpropagation.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def P():
    import photonprop
    return photonprop


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_lib(P):
    lib = P._lib.load_library()
    assert lib.pp_device_count() >= 1, "GPU test run without a visible HIP device"
    return lib


@pytest.fixture()
def dev(P, gpu_lib):
    """A fresh device context (capacity 2^20 hits, deterministic order)."""
    _, d = P.make_hit_buffers(capacity=1 << 20)
    yield d
    d.close()
