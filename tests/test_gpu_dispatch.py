"""Calls larger than one dispatch's 2^32 - 1 work-items (2^24 - 8 workgroups
of 256): 180,000 reflections x 25,323 voxels at a requested tile of 256
voxels would be 99 tiles x 180,000 = 17.8M workgroups.  The trace runs such a
call as several launches of whole reflections at the requested tile, so its
results equal those of a call that fits one launch bit for bit, in any launch
order; the bisection path re-tiles the call to fit one dispatch instead
(another FP32 summation order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_BLOCKS = (1 << 24) - 8


def test_oversized_call(gpu):
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    wl = W.single_gpu_config(R=180_000)
    sub = wl.subset(1500)
    print("case: 256^3 stride-2, 180,000 reflections, tile 256", flush=True)
    with Tracer.for_workload(wl) as t:
        t.set_tile(256)
        A = t.run_workload(wl)
        tiles = -(-len(wl.coords) // 256)
        assert t.stats()["workgroups"] == 180_000 * tiles > MAX_BLOCKS
        # launch groups of 64 reflections across the launch boundaries: each
        # launch groups its own reflections, the last group of each partly filled
        t.set_order(64, 1)
        Ag = t.run_workload(wl)
        t.set_order(0, -1)
        Ab = t.bisection(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F, 3)
        assert t.stats()["workgroups"] <= MAX_BLOCKS
        B = t.run_workload(sub)
        Bb = t.bisection(sub.s1, sub.omega, sub.raw_xray, sub.omega_axis, sub.F, 3)
    for a in (A, Ab):
        assert np.all(np.isfinite(a)) and np.all((a > 0) & (a <= 1))
    # the same reflections at the same tile in one launch: the same partial sums
    np.testing.assert_array_equal(A[:len(B)].view(np.uint64), B.view(np.uint64))
    np.testing.assert_array_equal(Ag.view(np.uint64), A.view(np.uint64))
    # the re-tiled bisection: another FP32 summation order
    np.testing.assert_allclose(Ab[:len(Bb)], Bb, rtol=1e-6, atol=0)
