"""A launch whose tile would need more than one dispatch's 2^32 - 1
work-items (2^24 - 8 workgroups of 256) is re-tiled to fit, so every
reflection is traced: 180,000 reflections x 25,323 voxels at a requested
tile of 256 voxels would be 99 tiles x 180,000 = 17.8M workgroups."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_BLOCKS = (1 << 24) - 8


def test_oversized_launch_is_retiled(gpu):
    from anacor_public_amd import workload as W
    from anacor_public_amd.ray_tracing import Tracer
    wl = W.single_gpu_config(R=180_000)
    sub = wl.subset(1500)
    print("case: 256^3 stride-2, 180,000 reflections, tile 256", flush=True)
    with Tracer.for_workload(wl) as t:
        t.set_tile(256)
        A = t.run_workload(wl)
        assert t.stats()["workgroups"] <= MAX_BLOCKS
        Ab = t.bisection(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F, 3)
        assert t.stats()["workgroups"] <= MAX_BLOCKS
        B = t.run_workload(sub)
        Bb = t.bisection(sub.s1, sub.omega, sub.raw_xray, sub.omega_axis, sub.F, 3)
    for a in (A, Ab):
        assert np.all(np.isfinite(a)) and np.all((a > 0) & (a <= 1))
    # the same reflections traced in a launch that fits at the requested tile
    # (another tiling, so another FP32 summation order)
    np.testing.assert_allclose(A[:len(B)], B, rtol=1e-6, atol=0)
    np.testing.assert_allclose(Ab[:len(Bb)], Bb, rtol=1e-6, atol=0)
