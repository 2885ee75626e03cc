"""Shared test setup: paths, the `gpu` marker, device fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dagger.jl_amd"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import dagger_mi355x._lib as L
        import ctypes as C
        n = C.c_int(0)
        L.load().dm_device_count(C.byref(n))
        return n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def dm():
    """The product package, bound to device 0.  GPU tests fail (not skip)
    when the native library does not load."""
    import dagger_mi355x
    emu = os.environ.get("DM_TEST_EMU_LIB")
    if emu:  # CPU emulation build (tests/cpu_emu), for host-side bounds checks only
        dagger_mi355x._lib.load(emu)
    dagger_mi355x.load_library()
    dagger_mi355x.init(0)
    return dagger_mi355x


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
