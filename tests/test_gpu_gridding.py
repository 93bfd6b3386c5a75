"""GPU tests of the gridding path: the absorption map built in one launch
equals the reference's ray_tracing_single_gridding (fixture) bit for bit as
float, the drop-in single-ray entry equals it as double, and the look-ups
(resident map and the drop-in nearest_neighbor_interpolate_overall on the
reference's padded layout) match the NumPy look-up oracle to 1e-12."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def golden_grid():
    from conftest import ROOT
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_gridding.npz"))


def _wl(g, R=24, seed=77):
    from anacor_public_amd import workload as W
    s1, om = W.make_reflections(R, seed)
    wl = W.Workload("golden-grid", g["volume"], g["coords"], s1, om)
    wl.voxel_size = g["voxel_size"]
    wl.coefficients = g["coefficients"]
    wl.full_iteration = int(g["full_iteration"])
    return wl


@pytest.mark.parametrize("layout", [0, 2])
def test_map_matches_reference(gpu, golden_grid, layout):
    from anacor_public_amd.gridding import Gridding
    from anacor_public_amd.ray_tracing import Tracer
    g = golden_grid
    wl = _wl(g)
    say(f"case: 24x12 map of 48 voxels, layout {layout}")
    with Tracer.for_workload(wl) as t:
        t.set_layout(layout)
        with Gridding(t, int(g["n_theta"]), int(g["n_phi"])) as grid:
            np.testing.assert_array_equal(grid.directions, g["dirs"])
            np.testing.assert_array_equal(grid.angles, g["angles"])
            m = grid.map()
        ok = g["ok"]
        np.testing.assert_array_equal(m[ok], g["values"][ok].astype(np.float32))
        with Gridding(t, int(g["n_theta"]), int(g["n_phi"]), is_exp=True) as grid:
            me = grid.map()
        np.testing.assert_array_equal(me[g["exp_dirs"]], g["values_exp"].astype(np.float32))


def test_single_ray_drop_in(gpu, golden_grid):
    from anacor_public_amd import _lib
    from anacor_public_amd.ray_tracing import LabelTable
    g = golden_grid
    lib = _lib.load()
    table = LabelTable(g["volume"])
    shape = np.array(g["volume"].shape, dtype=np.int64)
    vs = np.ascontiguousarray(g["voxel_size"], dtype=np.float64)
    mu = np.ascontiguousarray(g["coefficients"], dtype=np.float64)
    say("case: ray_tracing_single_gridding drop-in")
    for d in (0, 37, 150, 287):
        for k in (0, 21, 47):
            c = np.ascontiguousarray(g["coords"][k], dtype=np.int64)
            v = np.ascontiguousarray(g["dirs"][d])
            r = lib.ray_tracing_single_gridding(c.ctypes.data, v.ctypes.data,
                                                float(g["angles"][d, 0]), float(g["angles"][d, 1]),
                                                vs.ctypes.data, mu.ctypes.data, table.ptr,
                                                shape.ctypes.data, int(g["full_iteration"]), d, 0)
            assert r == g["values"][d, k]
    # angles that do not belong to the vector: the reference asserts, here NaN
    c = np.ascontiguousarray(g["coords"][0], dtype=np.int64)
    v = np.ascontiguousarray(g["dirs"][5])
    r = lib.ray_tracing_single_gridding(c.ctypes.data, v.ctypes.data, 0.123, 0.456, vs.ctypes.data,
                                        mu.ctypes.data, table.ptr, shape.ctypes.data, 0, 5, 0)
    assert np.isnan(r)


@pytest.mark.parametrize("method", [1, 2])
def test_lookups_match_oracle(gpu, oracle, golden_grid, method):
    from anacor_public_amd import gridding as G
    from anacor_public_amd.ray_tracing import Tracer
    from oracle import gridding_oracle as GO
    g = golden_grid
    wl = _wl(g, R=40)
    say(f"case: look-ups, method {method}")
    with Tracer.for_workload(wl) as t, G.Gridding(t, int(g["n_theta"]), int(g["n_phi"])) as grid:
        m = grid.map()
        A = grid.run(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F, method=method)
    padded = G.pad_map(m, int(g["n_theta"]), int(g["n_phi"]))
    geo = oracle.reflection_setup(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F)
    ang = np.stack([geo["theta"][:, 0], geo["phi"][:, 0], geo["theta"][:, 1], geo["phi"][:, 1]],
                   axis=1)
    ref = GO.lookup(padded, grid.theta_list, grid.phi_list, ang, method)
    np.testing.assert_allclose(A, ref, rtol=1e-12, atol=0)
    drop = G.nearest_neighbor_interpolate_overall(wl.coords, wl.s1, wl.omega, wl.raw_xray,
                                                  wl.omega_axis, wl.F, wl.voxel_size,
                                                  wl.coefficients, grid.theta_list, grid.phi_list,
                                                  padded, method=method)
    np.testing.assert_allclose(drop, ref, rtol=1e-12, atol=0)
