import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def orc():
    import _libs
    L = _libs.oracle()
    if L is None:
        pytest.fail("oracle/liboracle.so is missing: run __graft_entry__.build() (or make -C oracle)")
    return L


@pytest.fixture(scope="session")
def ref():
    import _libs
    L = _libs.reference()
    if L is None:
        pytest.skip("compiled reference (oracle/_ref) not available")
    return L


@pytest.fixture(scope="session")
def emu():
    import _libs
    L = _libs.hostemu()
    if L is None:
        pytest.fail("tests/hostemu/libhostemu.so is missing: run __graft_entry__.build()")
    return L


@pytest.fixture(scope="session")
def ctx():
    """mcrt context on GPU 0 (GPU tests only)."""
    import mcrt_amd
    c = mcrt_amd.Context(0)
    yield c

GOLDEN = os.path.join(TESTS, "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)
