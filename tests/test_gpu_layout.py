"""GPU tests of the bricked volume layouts: every layout must give the same
labels along every ray, so absorption factors are bit-identical to the raw
layout's (same tiles, same summation order), in both the compensated FP32 and
the FP64 mean, with labels outside 0..3 (4-bit codes), irregular volume sizes
(bricks padded past the volume) and both full_iteration modes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def run_layouts(wl, layouts, f64=False, tile=512):
    from anacor_public_amd import _lib
    from anacor_public_amd.ray_tracing import Tracer
    out = {}
    with Tracer.for_workload(wl) as t:
        t.set_tile(tile)
        for lay in layouts:
            t.set_layout(lay)
            out[lay] = t.run_workload(wl, f64_mean=f64)
            expect = {_lib.LAYOUT_RAW: _lib.LAYOUT_RAW, _lib.LAYOUT_BRICK2: _lib.LAYOUT_BRICK2,
                      _lib.LAYOUT_BRICK4: _lib.LAYOUT_BRICK4}.get(lay)
            if expect is not None:
                assert t.layout == expect
    return out


@pytest.mark.parametrize("f64", [False, True])
def test_small_volume_all_layouts(gpu, f64):
    from anacor_public_amd import _lib
    from anacor_public_amd import workload as W
    wl = W.small_config(R=64)
    say(f"case: 64^3 layouts raw / brick2 / brick4 / auto-brick, f64={f64}")
    L = (_lib.LAYOUT_RAW, _lib.LAYOUT_BRICK2, _lib.LAYOUT_BRICK4, _lib.LAYOUT_BRICK)
    out = run_layouts(wl, L, f64=f64)
    for lay in L[1:]:
        np.testing.assert_array_equal(bits(out[lay]), bits(out[_lib.LAYOUT_RAW]),
                                      err_msg=f"layout {lay}")


def test_fuzz_cases_bricked(gpu, golden_fuzz):
    """48 fuzz volumes (irregular sizes, labels 0..4 and 7, both
    full_iteration modes, rotated goniometers): 4-bit bricks equal raw."""
    from anacor_public_amd import _lib
    from anacor_public_amd.ray_tracing import Tracer
    from conftest import workload_from
    n = int(golden_fuzz["n_cases"])
    for i in range(n):
        wl = workload_from(golden_fuzz, f"c{i}_")
        say(f"case: fuzz {i} volume {wl.volume.shape} full {wl.full_iteration}")
        out = run_layouts(wl, (_lib.LAYOUT_RAW, _lib.LAYOUT_BRICK), f64=True)
        np.testing.assert_array_equal(bits(out[_lib.LAYOUT_BRICK]), bits(out[_lib.LAYOUT_RAW]),
                                      err_msg=f"case {i}")
        if np.any(np.asarray(wl.volume) > 3):
            with Tracer.for_workload(wl) as t, pytest.raises(_lib.AnacorError):
                t.set_layout(_lib.LAYOUT_BRICK2)


def test_two_bit_bricks_irregular_volume(gpu, golden_fuzz):
    """Labels folded into 0..3 on the fuzz volumes: 2-bit bricks equal raw."""
    import dataclasses
    from anacor_public_amd import _lib
    from conftest import workload_from
    for i in (3, 17, 30, 44):
        wl = workload_from(golden_fuzz, f"c{i}_")
        vol = np.asarray(wl.volume).astype(np.int64)
        wl = dataclasses.replace(wl, volume=np.where(vol > 3, vol % 4, vol).astype(np.uint8))
        say(f"case: fuzz {i} folded to 0..3, volume {wl.volume.shape}")
        out = run_layouts(wl, (_lib.LAYOUT_RAW, _lib.LAYOUT_BRICK2), f64=False)
        np.testing.assert_array_equal(bits(out[_lib.LAYOUT_BRICK2]), bits(out[_lib.LAYOUT_RAW]))
