"""bench.py's host-side contract, on the CPU: every per-step traffic figure it
reports is one that a committed rocprofv3 summary under profiles/ states for
the same work shape, a shape without one reports null, and a request for more
than one GPU ends with a message instead of a measurement."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

LAYOUT = {"raw": 0, "brick2": 2, "brick4": 4}


def shape_rows():
    """(layout, block, tile, group) -> B/step from the work-shape table."""
    text = open(os.path.join(ROOT, "profiles",
                             "trace_kernel_work-shapes_plain-256cube.md")).read()
    rows = {}
    for m in re.finditer(r"^\| (raw|brick2|brick4) \| \((\d+), (\d+), (\d+)\) \| (auto|\d+) \| "
                         r"(\d+) \| [\d,]+ \| [\d,]+ \| [\d,]+ \| ([\d.]+) \|", text, re.M):
        tile = 0 if m.group(5) == "auto" else int(m.group(5))
        key = (LAYOUT[m.group(1)], tuple(int(m.group(i)) for i in (2, 3, 4)), tile,
               int(m.group(6)))
        rows[key] = float(m.group(7))
    return rows


def test_every_traffic_figure_is_on_record():
    rows = shape_rows()
    assert len(rows) == 36
    head = {k[1:]: v for k, v in bench.TRAFFIC_BYTES_PER_STEP.items()
            if k[0] == "plain-256^3-stride2"}
    assert set(head) == set(rows)
    for k, v in head.items():
        assert v == pytest.approx(rows[k], abs=1e-4), k
    large = [v for k, v in bench.TRAFFIC_BYTES_PER_STEP.items()
             if k[0] == "plain-1000x1000x800-100k-first10000"]
    text = open(os.path.join(ROOT, "profiles",
                             "trace_kernel_brick2_large-1000x1000x800.md")).read()
    m = re.search(r"([\d.]+) B per voxel step", text)
    assert len(large) == 1 and m and large[0] == pytest.approx(float(m.group(1)), abs=1e-3)


def test_traffic_lookup():
    steps = 101_723_984_094
    assert bench.traffic_for("plain-256^3-stride2", 0, [1, 2, 64], 0, 1, steps) == \
        pytest.approx(bench.TRAFFIC_BYTES_PER_STEP[("plain-256^3-stride2", 0, (1, 2, 64), 0, 1)]
                      * steps)
    # a tile, group or workload without a measured figure reports null
    assert bench.traffic_for("plain-256^3-stride2", 0, [1, 2, 64], 1024, 1, steps) is None
    assert bench.traffic_for("plain-256^3-stride2", 0, [1, 2, 64], 0, 2048, steps) is None
    assert bench.traffic_for("plain-64^3-all", 0, [1, 2, 64], 0, 1, steps) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
