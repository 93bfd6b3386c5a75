"""Pin the FP64 oracle against golden vectors written by the reference's own C
build (tests/golden/make_golden.py drives oracle/_ref/ray_tracing_cpu.so).

Everything here is bit-exact: A_h per reflection, the rotated vectors and beam
angles, and per ray the exit face, len_path_2, last path coordinate and the
four per-material lengths.
"""
import dataclasses

import numpy as np
import pytest

from conftest import workload_from


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_small_config_A_h_bit_exact(oracle, golden_small):
    wl = workload_from(golden_small)
    assert int(golden_small["vol_sum"]) > 0
    out = oracle.trace(wl, n_threads=4)
    assert out["failed"] == 0
    np.testing.assert_array_equal(bits(out["result"]), bits(golden_small["A_full0"]))


def test_small_config_full_iteration_bit_exact(oracle, golden_small):
    wl = dataclasses.replace(workload_from(golden_small), full_iteration=1)
    out = oracle.trace(wl, n_threads=4)
    np.testing.assert_array_equal(bits(out["result"]), bits(golden_small["A_full1"]))


def test_small_config_geometry_bits(oracle, golden_small):
    wl = workload_from(golden_small)
    g = oracle.reflection_setup(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F)
    for k in ("xray", "s1", "theta", "phi", "dir"):
        np.testing.assert_array_equal(bits(g[k]), bits(golden_small[f"geom_{k}"]), err_msg=k)


def _check_rays(rays, face, length, last, L, seg_off):
    np.testing.assert_array_equal(rays["face"], face)
    np.testing.assert_array_equal(rays["len"], length)
    multi = length > 1
    np.testing.assert_array_equal(rays["last"][multi], last[multi])
    np.testing.assert_array_equal(bits(rays["L"]), bits(L))
    nseg = np.diff(seg_off).reshape(face.shape)
    np.testing.assert_array_equal(rays["nseg"], nseg)


def test_small_config_per_ray_records(oracle, golden_small):
    wl = workload_from(golden_small)
    vidx = golden_small["ray_voxels"]
    R = golden_small["ray_face"].shape[0]
    sub = dataclasses.replace(wl, coords=wl.coords[vidx], s1=wl.s1[:R], omega=wl.omega[:R])
    out = oracle.trace(sub, n_threads=4, rays=True)
    _check_rays(out["rays"], golden_small["ray_face"], golden_small["ray_len"],
                golden_small["ray_last"], golden_small["ray_L"], golden_small["ray_seg_off"])


def _fuzz_cases(golden_fuzz):
    return range(int(golden_fuzz["n_cases"]))


def test_fuzz_A_h_and_abort_agreement(oracle, golden_fuzz):
    n_ok = n_abort = 0
    for c in _fuzz_cases(golden_fuzz):
        wl = workload_from(golden_fuzz, f"c{c}_")
        out = oracle.trace(wl, n_threads=1)
        ok = golden_fuzz[f"c{c}_ok"]
        ref = golden_fuzz[f"c{c}_A"]
        res = out["result"]
        # reflections on which the reference aborts are exactly the oracle's NaNs
        np.testing.assert_array_equal(np.isnan(res), ~ok, err_msg=f"case {c}")
        np.testing.assert_array_equal(bits(res[ok]), bits(ref[ok]), err_msg=f"case {c}")
        n_ok += int(ok.sum())
        n_abort += int((~ok).sum())
    assert n_ok > 200 and n_abort > 5


def test_fuzz_per_ray_records(oracle, golden_fuzz):
    checked = 0
    for c in _fuzz_cases(golden_fuzz):
        key = f"c{c}_good"
        if key not in golden_fuzz:
            continue
        wl = workload_from(golden_fuzz, f"c{c}_")
        good = golden_fuzz[key]
        sub = dataclasses.replace(wl, s1=wl.s1[good], omega=wl.omega[good])
        out = oracle.trace(sub, n_threads=1, rays=True)
        _check_rays(out["rays"], golden_fuzz[f"c{c}_ray_face"], golden_fuzz[f"c{c}_ray_len"],
                    golden_fuzz[f"c{c}_ray_last"], golden_fuzz[f"c{c}_ray_L"],
                    golden_fuzz[f"c{c}_ray_seg_off"])
        g = oracle.reflection_setup(sub.s1, sub.omega, sub.raw_xray, sub.omega_axis, sub.F)
        for k in ("theta", "phi", "dir", "xray", "s1"):
            np.testing.assert_array_equal(bits(g[k]), bits(golden_fuzz[f"c{c}_geom_{k}"]))
        checked += out["rays"].size
    assert checked > 1000
