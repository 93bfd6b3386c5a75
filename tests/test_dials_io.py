"""The DIALS-derived wire formats (CPU): the column parser reads the same
numbers as the reference's per-row literal_eval, the goniometer follows
main_lite.py:211-225, corrections round-trip and stack in chunk order."""
import json
from ast import literal_eval

import numpy as np


def _write_refl(path, s1, xyz, rng):
    rows = []
    for v, o in zip(s1, xyz):
        rows.append({"s1": str(tuple(float(x) for x in v)),
                     "xyzobs.mm.value": str(tuple(float(x) for x in o)),
                     "intensity.sum.value": str(float(rng.normal())),
                     "miller_index": str(tuple(int(k) for k in rng.integers(-20, 20, 3)))})
    with open(path, "w") as f:
        json.dump(rows, f)
    return rows


def test_reflections_parse_like_literal_eval(tmp_path):
    from anacor_public_amd import dials_io, workload as W
    rng = np.random.default_rng(0)
    s1, om = W.make_reflections(500, 3)
    xyz = np.stack([rng.uniform(0, 2000, 500), rng.uniform(0, 2000, 500), om], axis=1)
    xyz[7, 2] = 1e-320                       # a subnormal and an exponent form survive
    rows = _write_refl(tmp_path / "r.json", s1, xyz, rng)
    got = dials_io.load_reflections(str(tmp_path / "r.json"), low=10, up=400, offset_deg=12.5)
    ref_s1 = np.array([literal_eval(r["s1"]) for r in rows[10:400]])
    ref_om = np.array([literal_eval(r["xyzobs.mm.value"])[2] for r in rows[10:400]])
    ref_om = ref_om + 12.5 / 180 * np.pi
    np.testing.assert_array_equal(got.s1, ref_s1)
    np.testing.assert_array_equal(got.omega, ref_om)
    np.testing.assert_array_equal(got.miller_index[3], literal_eval(rows[13]["miller_index"]))
    tail = dials_io.load_reflections(str(tmp_path / "r.json"), low=450, up=-1)
    assert len(tail.omega) == 50


def test_experiment_goniometer(tmp_path):
    from anacor_public_amd import dials_io
    from anacor_public_amd.ray_tracing import kp_rotation
    kappa = np.array([0.914, 0.279, -0.297])
    kappa = (kappa / np.linalg.norm(kappa)).tolist()
    gon = {"axes": [[1.0, 0.0, 0.0], kappa, [1.0, 0.0, 0.0]],
           "angles": [30.0, -45.0, 0.0], "names": ["PHI", "KAPPA", "OMEGA"], "scan_axis": 2}
    beam = {"direction": [0.0, 0.0, 1.0], "wavelength": 0.9}
    with open(tmp_path / "e.json", "w") as f:
        json.dump([gon, beam], f)
    e = dials_io.load_experiment(str(tmp_path / "e.json"))
    np.testing.assert_array_equal(e.raw_xray, [-0.0, -0.0, -1.0])
    np.testing.assert_array_equal(e.omega_axis, [1.0, 0.0, 0.0])
    F = kp_rotation(np.array(gon["axes"][1]), -45.0 / 180 * np.pi) @ \
        kp_rotation(np.array(gon["axes"][0]), 30.0 / 180 * np.pi)
    np.testing.assert_allclose(e.F.reshape(3, 3), F, rtol=0, atol=1e-15)
    # a rotation matrix
    np.testing.assert_allclose(e.F.reshape(3, 3) @ e.F.reshape(3, 3).T, np.eye(3), atol=1e-14)


def test_corrections_round_trip_and_stacking(tmp_path):
    from anacor_public_amd import dials_io
    chunks = {100: [0.5, 0.25], 2: [0.125], 30: [0.75, 0.875, 0.9], -1: [0.1]}
    for up, c in chunks.items():
        dials_io.save_corrections(str(tmp_path), "16010", up, c)
    with open(tmp_path / "16010_refl_30.json") as f:
        assert json.load(f) == [0.75, 0.875, 0.9]
    out = dials_io.stack_corrections(str(tmp_path), "16010")
    assert out == [0.125, 0.75, 0.875, 0.9, 0.5, 0.25, 0.1]
    with open(tmp_path / "16010_refl_overall.json") as f:
        assert json.load(f) == out


import pytest  # noqa: E402


@pytest.mark.gpu
def test_files_end_to_end(gpu, oracle, tmp_path):
    """Reflection and experiment JSON in, the tracer, chunked correction files
    out and stacked: the same numbers as tracing the arrays directly, within
    1e-6 of the FP64 oracle."""
    from anacor_public_amd import dials_io, workload as W
    from anacor_public_amd.ray_tracing import Tracer
    wl = W.small_config(R=40)
    rng = np.random.default_rng(1)
    xyz = np.stack([rng.uniform(0, 100, 40), rng.uniform(0, 100, 40), wl.omega], axis=1)
    # the stacker takes every *refl*.json of its folder, as the reference's
    # does, so the inputs live apart from the correction chunks
    inp, out_dir = tmp_path / "in", tmp_path / "out"
    inp.mkdir()
    out_dir.mkdir()
    _write_refl(inp / "refl.json", wl.s1, xyz, rng)
    gon = {"axes": [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], wl.omega_axis.tolist()],
           "angles": [0.0, 0.0, 0.0]}
    with open(inp / "expt.json", "w") as f:
        json.dump([gon, {"direction": (-wl.raw_xray).tolist()}], f)
    e = dials_io.load_experiment(str(inp / "expt.json"))
    np.testing.assert_array_equal(e.F, np.eye(3).reshape(-1))
    with Tracer.for_workload(wl) as t:
        direct = t.run_workload(wl)
        for low, up in ((0, 20), (20, -1)):
            r = dials_io.load_reflections(str(inp / "refl.json"), low=low, up=up)
            dials_io.save_corrections(str(out_dir), "ds", up,
                                      t.run(r.s1, r.omega, e.raw_xray, e.omega_axis, e.F))
    out = np.array(dials_io.stack_corrections(str(out_dir), "ds"))
    np.testing.assert_array_equal(out, direct)
    np.testing.assert_allclose(out, oracle.trace(wl)["result"], rtol=1e-6, atol=0)
