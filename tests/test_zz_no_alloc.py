"""The allocation-free trace entry point (anacor_ctx_run_noalloc) against the
entry point it mirrors (anacor_ctx_run), called through the C-ABI.

Both use the same tile wherever reflections x tiles fit the static
partial-sum capacity, so their means must agree bit for bit; two calls in a
row, the second with fewer reflections, check that nothing from the first
call (error words, step counter, partial sums) leaks into the second."""
import ctypes as ct

import numpy as np
import pytest

from anacor_public_amd import _lib
from anacor_public_amd import workload as W
from anacor_public_amd.ray_tracing import Tracer

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _call(fn, t, wl, R, flags=0):
    s1 = np.ascontiguousarray(wl.s1[:R], dtype=np.float64)
    om = np.ascontiguousarray(wl.omega[:R], dtype=np.float64)
    out = np.full(max(R, 1), -7.0)
    rc = fn(t.handle, _p(s1), _p(om), R, _p(np.ascontiguousarray(wl.raw_xray, np.float64)),
            _p(np.ascontiguousarray(wl.omega_axis, np.float64)),
            _p(np.ascontiguousarray(wl.F, np.float64)), flags, _p(out))
    st = t.stats()
    return rc, out[:R], st["voxel_steps"]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("flags", [0, _lib.FLAG_F64_MEAN])
def test_two_calls_match_ctx_run(gpu, flags):
    wl = W.small_config(R=256)
    with Tracer.for_workload(wl) as t:
        rc1, na1, st1 = _call(gpu.anacor_ctx_run_noalloc, t, wl, 256, flags)
        rc2, na2, st2 = _call(gpu.anacor_ctx_run_noalloc, t, wl, 100, flags)
        rr1, ref1, sr1 = _call(gpu.anacor_ctx_run, t, wl, 256, flags)
        rr2, ref2, sr2 = _call(gpu.anacor_ctx_run, t, wl, 100, flags)
    assert (rc1, rc2, rr1, rr2) == (0, 0, 0, 0)
    np.testing.assert_array_equal(_bits(na1), _bits(ref1))
    np.testing.assert_array_equal(_bits(na2), _bits(ref2))
    assert (st1, st2) == (sr1, sr2)
    assert np.all((na2 > 0) & (na2 <= 1))


def test_bricked_layout_and_largest_call(gpu):
    """The 256^3 workload at the maximum reflection count in 2-bit bricks
    (built beforehand by anacor_ctx_set_layout), then a smaller raw call."""
    wl = W.single_gpu_config(R=_lib.NOALLOC_MAX_REFLECTIONS)
    R = _lib.NOALLOC_MAX_REFLECTIONS
    with Tracer.for_workload(wl) as t:
        t.set_layout(_lib.LAYOUT_BRICK2)
        rc1, na1, _ = _call(gpu.anacor_ctx_run_noalloc, t, wl, R)
        rr1, ref1, _ = _call(gpu.anacor_ctx_run, t, wl, R)
        t.set_layout(_lib.LAYOUT_RAW)
        rc2, na2, _ = _call(gpu.anacor_ctx_run_noalloc, t, wl, 3000)
        rr2, ref2, _ = _call(gpu.anacor_ctx_run, t, wl, 3000)
    assert (rc1, rr1, rc2, rr2) == (0, 0, 0, 0)
    np.testing.assert_array_equal(_bits(na1), _bits(ref1))
    np.testing.assert_array_equal(_bits(na2), _bits(ref2))


def test_geometry_errors_as_ctx_run(gpu, golden_fuzz):
    """Reflections on which the reference aborts: NaN and ANACOR_ERR_GEOMETRY
    from both entry points alike."""
    from conftest import workload_from
    hit = 0
    for case in range(int(golden_fuzz["n_cases"])):
        wl = workload_from(golden_fuzz, f"c{case}_")
        R = len(wl.omega)
        with Tracer.for_workload(wl) as t:
            rc_na, na, _ = _call(gpu.anacor_ctx_run_noalloc, t, wl, R)
            rc_ref, ref, _ = _call(gpu.anacor_ctx_run, t, wl, R)
        assert rc_na == rc_ref
        np.testing.assert_array_equal(np.isnan(na), np.isnan(ref))
        ok = ~np.isnan(ref)
        np.testing.assert_array_equal(_bits(na[ok]), _bits(ref[ok]))
        hit += int(rc_ref == _lib.ANACOR_ERR_GEOMETRY)
    assert hit > 0


def test_rejects_more_than_the_maximum(gpu):
    wl = W.small_config(R=4)
    R = _lib.NOALLOC_MAX_REFLECTIONS + 1
    big = W.Workload("big", wl.volume, wl.coords, np.tile(wl.s1[:1], (R, 1)),
                     np.tile(wl.omega[:1], R))
    with Tracer.for_workload(wl) as t:
        rc, out, _ = _call(gpu.anacor_ctx_run_noalloc, t, big, R)
        code, msg = _lib.last_error(gpu)
        assert np.all(out == -7.0)   # nothing written
    assert rc == _lib.ANACOR_ERR_ARG and code == _lib.ANACOR_ERR_ARG
    assert "ANACOR_NOALLOC_MAX_REFLECTIONS" in msg
