import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_npz(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def workload_from(npz, prefix=""):
    from anacor_public_amd import workload as W
    g = lambda k: npz[prefix + k]  # noqa: E731
    wl = W.Workload("golden", g("volume"), g("coords"), g("s1"), g("omega"))
    wl.raw_xray = g("raw_xray")
    wl.omega_axis = g("omega_axis")
    wl.F = g("F")
    wl.voxel_size = g("voxel_size")
    wl.coefficients = g("coefficients")
    wl.full_iteration = int(g("full_iteration"))
    return wl


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import ORACLE_SO, Oracle, build
    if not os.path.exists(ORACLE_SO):
        build(with_reference=False)
    return Oracle()


@pytest.fixture(scope="session")
def golden_small():
    return load_npz("golden_small.npz")


@pytest.fixture(scope="session")
def golden_fuzz():
    return load_npz("golden_fuzz.npz")


def _gpu_available():
    try:
        from anacor_public_amd import _lib
        return _lib.load().anacor_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU gate: a gpu-marked test fails loudly when the HIP library
    or the device is missing."""
    from anacor_public_amd import _lib
    lib = _lib.load()
    assert lib.anacor_device_count() > 0, "no HIP device visible"
    return lib
