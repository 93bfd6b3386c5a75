"""Autotuner for the trace kernels' work shapes.

Three knobs change speed but never a path length:

* the wave grouping of crystal voxels (``(bz, by, bx)`` blocks, lanes along
  x; ``(0, 0, 0)`` = Morton order), which decides how many volume rows a
  wave's 64 parallel rays touch per step;
* the tile size (voxels per workgroup), which trades partial sums and
  scheduling granularity against load balance;
* the volume layout (raw uint8 or 8x8x8 bricks of 2-/4-bit labels) of a
  single-sample tracer (the batched kernel reads the raw layout).

``tune(tracer, ...)`` times every candidate on a slice of the sample's own
reflections with the trace kernel's HIP-event duration and keeps the fastest.
Results are cached per (volume shape, voxel count, voxel-lattice stride) in a
JSON file (``$ANACOR_TUNE_CACHE``, default ``~/.cache/anacor_mi355x/tune.json``).
"""
from __future__ import annotations

import json
import os
import threading
from typing import Dict, Iterable, Optional, Sequence, Tuple

import numpy as np

BLOCKS: Tuple[Tuple[int, int, int], ...] = ((1, 2, 64), (1, 1, 128), (2, 2, 32), (1, 4, 32),
                                             (2, 4, 16), (0, 0, 0))
TILES: Tuple[int, ...] = (0, 512, 2048)       # 0 = the library's automatic choice (1,024 max)
LAYOUTS: Tuple[int, ...] = (0, 1)             # raw, bricks (_lib.LAYOUT_RAW, _lib.LAYOUT_BRICK)

_lock = threading.Lock()


def cache_path() -> str:
    return os.environ.get("ANACOR_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "anacor_mi355x",
                                       "tune.json"))


def _lattice_stride(coords: np.ndarray) -> int:
    """Largest s with all voxels on one lattice of spacing s (1 for all voxels)."""
    c = np.asarray(coords).reshape(-1, 3)
    if len(c) < 2:
        return 1
    g = 0
    for a in range(3):
        d = np.unique(c[:, a]) - c[:, a].min()
        g = int(np.gcd(g, np.gcd.reduce(d)))
    return max(g, 1)


def sample_key(shape, coords) -> str:
    return f"{tuple(int(s) for s in shape)}|N={len(np.asarray(coords).reshape(-1, 3))}|stride={_lattice_stride(coords)}"


def _load_cache() -> Dict[str, dict]:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, entry: dict) -> None:
    with _lock:
        data = _load_cache()
        data[key] = entry
        path = cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass


def tune(tracer, s1, omega, raw_xray, omega_axis, F, coords, trial: int = 512,
         blocks: Sequence[Tuple[int, int, int]] = BLOCKS, tiles: Iterable[int] = TILES,
         layouts: Optional[Sequence[int]] = None, use_cache: bool = True,
         repeats: int = 2) -> dict:
    """Pick the fastest (layout, block, tile) for ``tracer`` (a Tracer, or a
    borrowed batch-sample handle with ``layouts=None``) on its first ``trial``
    reflections, apply it, and return {"layout", "block", "tile", "ms",
    "table"}.  layouts None leaves the layout alone."""
    key = sample_key(tracer.shape, coords) + ("" if layouts is None else "|layouts")
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            if layouts is not None:
                tracer.set_layout(int(hit["layout"]))
            tracer.set_voxel_block(*hit["block"])
            tracer.set_tile(int(hit["tile"]))
            return dict(hit, cached=True)
    s1 = np.ascontiguousarray(s1)[:trial]
    omega = np.ascontiguousarray(omega)[:trial]
    table = []
    best = None
    for lay in (layouts if layouts is not None else (None,)):
        if lay is not None:
            tracer.set_layout(int(lay))
        for blk in blocks:
            tracer.set_voxel_block(*blk)
            for tile in tiles:
                tracer.set_tile(int(tile))
                tracer.run(s1, omega, raw_xray, omega_axis, F)          # warm
                ms = min(_timed(tracer, s1, omega, raw_xray, omega_axis, F)
                         for _ in range(repeats))
                row = {"layout": lay, "block": list(blk), "tile": int(tile), "ms": ms}
                table.append(row)
                if best is None or ms < best["ms"]:
                    best = dict(row)
    if best["layout"] is not None:
        tracer.set_layout(int(best["layout"]))
    tracer.set_voxel_block(*best["block"])
    tracer.set_tile(best["tile"])
    entry = dict(best, table=table, trial=int(len(omega)))
    if use_cache:
        _store_cache(key, entry)
    return dict(entry, cached=False)


def _timed(tracer, s1, omega, raw_xray, omega_axis, F) -> float:
    tracer.run(s1, omega, raw_xray, omega_axis, F)
    return float(tracer.stats()["kernel_ms"])


def tune_batch(batch, workloads, trial: int = 512, **kw) -> list:
    """Tune every sample of a Batch on its own reflections."""
    out = []
    for i, wl in enumerate(workloads):
        out.append(tune(batch.sample(i), wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F,
                        wl.coords, trial=trial, **kw))
    return out
