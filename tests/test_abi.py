"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-side FP64 geometry (rotation, beam angles,
per-face deltas) carries the same bits as the oracle."""
import ctypes as ct
import dataclasses
import os
import re

import numpy as np
import pytest

from conftest import ROOT, workload_from


def header_symbols():
    text = open(os.path.join(ROOT, "include", "anacor_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "defined", "sizeof", "extern"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_header_and_exports_agree():
    from anacor_public_amd import _lib
    declared = header_symbols()
    assert set(declared) == set(_lib.EXPORTED), (declared, _lib.EXPORTED)
    lib = ct.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_drop_in_aliases_export_reference_names():
    from anacor_public_amd import build
    for alias in build.ALIASES:
        lib = ct.CDLL(os.path.join(build.LIBDIR, alias))
        for name in ("ray_tracing_overall", "ray_tracing_gpu_overall", "ray_tracing_single",
                     "ray_tracing_single_mp", "free"):
            assert hasattr(lib, name)


def test_device_count_without_gpu_is_nonnegative():
    from anacor_public_amd import _lib
    assert _lib.load().anacor_device_count() >= 0


def test_gather_probe_rejects_bad_arguments():
    from anacor_public_amd import _lib
    lib = _lib.load()
    rate = ct.c_double(-1.0)
    for k, pattern, loads in ((31, 0, 1 << 20), (5, 0, 1 << 20), (24, 2, 1 << 20), (24, 1, 0)):
        assert lib.anacor_probe_gather(0, k, pattern, loads, ct.byref(rate)) == _lib.ANACOR_ERR_ARG
    assert lib.anacor_probe_gather(0, 24, 0, 1 << 20, None) == _lib.ANACOR_ERR_ARG


def _bits(a):
    """Bit patterns, every NaN mapped to one: IEEE 754 fixes no sign or payload
    for a NaN that arithmetic produces, and compilers may legally move a
    negation across it (-(a/b) as (-a)/b), as the clang host build does on
    the delta-table entries of faces a ray cannot leave through."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    return np.where(np.isnan(a), np.nan, a).view(np.uint64)


def _oracle_d_table(oracle, g):
    R = len(g)
    tab = np.zeros((R, 2, 6, 3))
    for i in range(R):
        for k in range(2):
            for f in range(1, 7):
                tab[i, k, f - 1] = oracle.face_rule(g["theta"][i, k], g["phi"][i, k], f)["d"]
    return tab


@pytest.mark.parametrize("case", ["small", "fuzz"])
def test_host_geometry_matches_oracle_bits(oracle, golden_small, golden_fuzz, case):
    from anacor_public_amd.ray_tracing import reflection_geometry
    wls = [workload_from(golden_small)] if case == "small" else \
        [workload_from(golden_fuzz, f"c{c}_") for c in range(int(golden_fuzz["n_cases"]))]
    for wl in wls:
        tp, dt = reflection_geometry(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F,
                                     int(max(wl.volume.shape)))
        g = oracle.reflection_setup(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F)
        np.testing.assert_array_equal(_bits(tp[:, [0, 2]]), _bits(g["theta"]))
        np.testing.assert_array_equal(_bits(tp[:, [1, 3]]), _bits(g["phi"]))
        np.testing.assert_array_equal(_bits(dt), _bits(_oracle_d_table(oracle, g)))


def test_host_geometry_matches_reference_goldens(golden_small):
    from anacor_public_amd.ray_tracing import reflection_geometry
    wl = workload_from(golden_small)
    tp, _ = reflection_geometry(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F, 64)
    np.testing.assert_array_equal(_bits(tp[:, [0, 2]]), _bits(golden_small["geom_theta"]))
    np.testing.assert_array_equal(_bits(tp[:, [1, 3]]), _bits(golden_small["geom_phi"]))


def test_no_device_returns_nan_and_error_code():
    """Without a HIP device the drop-in entry returns a NaN-filled malloc'd
    array and reports ANACOR_ERR_NO_DEVICE instead of aborting."""
    from anacor_public_amd import _lib, ray_tracing as rt
    from anacor_public_amd import workload as W
    lib = _lib.load()
    if lib.anacor_device_count() > 0:
        pytest.skip("a HIP device is visible; this checks the device-less error path")
    vol = np.ones((8, 8, 8), np.uint8)
    coords = np.array([[4, 4, 4]], np.int64)
    s1, om = W.make_reflections(3, 1)
    with pytest.raises(_lib.AnacorError) as ei:
        rt.ray_tracing_overall(0, 3, coords, 1, s1, om, W.RAW_XRAY, W.OMEGA_AXIS, W.F_IDENTITY, 3,
                               W.VOXEL_SIZE, W.COEFFICIENTS, vol, np.array(vol.shape))
    assert ei.value.code == _lib.ANACOR_ERR_NO_DEVICE
