"""The data formats on either side of the ray-tracing path.

* Reflections: the JSON list that lite/refl_2_json.py:65-110 writes from a
  DIALS reflection table -- one dict per reflection, every value the str() of
  the DIALS column entry, e.g. ``"s1": "(0.0123, -0.998, 0.061)"``.  The
  reference parses every row with ``literal_eval`` (utils/utils_mp.py:151-164);
  ``load_reflections`` parses whole columns at once (same correctly rounded
  float conversion, so the same bits) and applies the rotation offset the same
  way.
* Experiment: ``[goniometer.to_dict(), beam.to_dict()]``
  (lite/refl_2_json.py:112-118); ``load_experiment`` derives the goniometer
  as main_lite.py:211-225 and utils_mp.py:48 do: F = R(kappa) R(phi) from the
  first two axes, the omega axis from the third, the beam ``raw_xray`` as minus
  the beam direction.
* Corrections: ``{dataset}_refl_{up}.json``, an indented JSON list of floats per
  worker chunk (utils_mp.py:426-427), merged by utils/stacking.py:40-105
  (chunks sorted by their number, the ``-1`` chunk last) into
  ``{dataset}_refl_overall.json``.
"""
from __future__ import annotations

import json
import os
import re
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .ray_tracing import kp_rotation


@dataclass
class Experiment:
    raw_xray: np.ndarray      # (3,) minus the beam direction
    omega_axis: np.ndarray    # (3,)
    F: np.ndarray             # (9,) R(kappa) . R(phi), row-major


@dataclass
class Reflections:
    s1: np.ndarray            # [R][3]
    omega: np.ndarray         # [R] rotation angle (rad) plus the offset
    intensity: Optional[np.ndarray]
    miller_index: Optional[np.ndarray]


def load_experiment(path: str) -> Experiment:
    with open(path) as f:
        axes_data = json.load(f)
    axes = axes_data[0]
    kappa_matrix = kp_rotation(np.array(axes["axes"][1]), axes["angles"][1] / 180 * np.pi)
    phi_matrix = kp_rotation(np.array(axes["axes"][0]), axes["angles"][0] / 180 * np.pi)
    F = np.dot(kappa_matrix, phi_matrix)
    return Experiment(raw_xray=-np.array(axes_data[1]["direction"], dtype=np.float64),
                      omega_axis=np.array(axes["axes"][2], dtype=np.float64),
                      F=np.ascontiguousarray(F, dtype=np.float64).reshape(-1))


_STRIP = re.compile(r"[()\[\]\s]")


def parse_tuples(column: Sequence[str], width: int) -> np.ndarray:
    """[n][width] floats from strings such as "(1.0, -2.5e-3, 3)"."""
    if len(column) == 0:
        return np.zeros((0, width))
    flat = _STRIP.sub("", ",".join(column)).split(",")
    out = np.array(flat, dtype=np.float64)
    if out.size != len(column) * width:
        raise ValueError(f"expected {width} numbers per entry")
    return out.reshape(len(column), width)


def load_reflections(path: str, low: int = 0, up: int = -1, offset_deg: float = 0.0,
                     data: Optional[list] = None) -> Reflections:
    """Reflections low..up (up = -1: to the end, main_lite.py:178-184) with
    omega = xyzobs.mm.value[2] + offset (utils_mp.py:151-164)."""
    if data is None:
        with open(path) as f:
            data = json.load(f)
    rows = data[low:] if up == -1 else data[low:up]
    s1 = parse_tuples([r["s1"] for r in rows], 3)
    xyz = parse_tuples([r["xyzobs.mm.value"] for r in rows], 3)
    omega = xyz[:, 2] + offset_deg / 180 * np.pi
    inten = (np.array([float(r["intensity.sum.value"]) for r in rows])
             if rows and "intensity.sum.value" in rows[0] else None)
    miller = (parse_tuples([r["miller_index"] for r in rows], 3).astype(np.int64)
              if rows and "miller_index" in rows[0] else None)
    return Reflections(np.ascontiguousarray(s1), np.ascontiguousarray(omega), inten, miller)


def save_corrections(save_dir: str, dataset: str, up: int, corr) -> str:
    path = os.path.join(save_dir, "{}_refl_{}.json".format(dataset, up))
    with open(path, "w") as f:
        json.dump([float(c) for c in np.asarray(corr).reshape(-1)], f, indent=2)
    return path


def _chunk_number(name: str) -> int:
    found = re.findall(r"(\d+)", name)
    return int(found[-1]) if found else -1


def stack_corrections(path: str, dataset: str, write: bool = True) -> List[float]:
    """Concatenate the correction chunk files in `path` in chunk order, the
    ``-1`` (last) chunk at the end, as utils/stacking.py:40-105 does (every
    ``*refl*.json`` chunk of the folder, whatever its dataset prefix)."""
    names, last = [], None
    for name in os.listdir(path):
        if "json" not in name or "overall" in name or "refl" not in name or "dict" in name:
            continue
        if "-1" in name:
            last = name
            continue
        names.append(name)
    names.sort(key=_chunk_number)
    if last:
        names.append(last)
    corr: List[float] = []
    for name in names:
        with open(os.path.join(path, name)) as f:
            corr += json.load(f)
    if write:
        with open(os.path.join(path, "{}_refl_overall.json".format(dataset)), "w+") as f:
            json.dump(corr, f, indent=2)
    return corr
