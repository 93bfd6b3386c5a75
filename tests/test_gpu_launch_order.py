"""GPU tests of the trace launch order: reflections in launch groups run tile
by tile, with either ray leading the direction sort.  The order only decides
which workgroup computes which (reflection, tile) partial; every partial has
its own slot and is merged in tile order, so every order must give the same
absorption factors bit for bit — including groups larger than the call, a
last group that is only partly filled, and both volume layouts at a fixed
tile."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("layout_name", ["RAW", "BRICK2"])
def test_orders_are_bit_identical(gpu, oracle, layout_name):
    from anacor_public_amd import _lib
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    wl = W.small_config(R=301)              # 3,169 voxels: 13 tiles of 256
    layout = getattr(_lib, f"LAYOUT_{layout_name}")
    with Tracer.for_workload(wl) as t:
        t.set_layout(layout)
        t.set_tile(256)
        t.set_order(1, 0)
        ref = t.run_workload(wl)
        assert t.work_shape()["group"] == 1 and t.work_shape()["direction_major"] == 0
        for group in (7, 64, 2048):          # 301 = 43 x 7; 4 x 64 + 45; one partial group
            for major in (0, 1):
                say(f"case: {layout_name}, group {group}, direction_major {major}")
                t.set_order(group, major)
                got = t.run_workload(wl)
                ws = t.work_shape()
                assert (ws["group"], ws["direction_major"]) == (group, major)
                np.testing.assert_array_equal(bits(got), bits(ref),
                                              err_msg=f"group {group} major {major}")
        t.set_order(0, -1)                   # automatic again
        ws = t.work_shape()
        bricked = layout != _lib.LAYOUT_RAW
        assert ws["group"] == (2048 if bricked else 1)
        assert ws["direction_major"] == (1 if bricked else 0)
        auto = t.run_workload(wl)
    np.testing.assert_array_equal(bits(auto), bits(ref))
    exp = oracle.trace(wl.subset(4))["result"]
    np.testing.assert_allclose(ref[:4], exp, rtol=1e-6, atol=0)


def test_order_argument_checks(gpu):
    from anacor_public_amd import _lib
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    wl = W.small_config(R=4)
    with Tracer.for_workload(wl) as t:
        for bad in ((-1, 0), (1, 2), (1, -2)):
            with pytest.raises(_lib.AnacorError):
                t.set_order(*bad)


@pytest.mark.parametrize("layout_name", ["RAW", "BRICK2"])
def test_two_part_call_equals_one_part_calls(gpu, layout_name):
    """A call of 8,192 reflections or more runs in two parts (the last three
    quarters set up on the host while the first quarter is traced); at a
    fixed tile its results equal, bit for bit, those of calls below that
    size, which run in one part."""
    from anacor_public_amd import _lib
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    wl = W.small_config(R=12001)            # parts of 3,000 and 9,001
    geo = (wl.raw_xray, wl.omega_axis, wl.F)
    say(f"case: two-part call, {layout_name}, R = {len(wl.omega)}")
    with Tracer.for_workload(wl) as t:
        t.set_layout(getattr(_lib, f"LAYOUT_{layout_name}"))
        t.set_tile(256)
        whole = t.run_workload(wl)
        parts = np.concatenate([t.run(wl.s1[i:i + 4000], wl.omega[i:i + 4000], *geo)
                                for i in range(0, len(wl.omega), 4000)])
    np.testing.assert_array_equal(bits(whole), bits(parts))
    assert np.mean(np.isfinite(whole)) > 0.99
