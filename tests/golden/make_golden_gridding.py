"""Write tests/golden/golden_gridding.npz from the reference's own C code.

The reference's ``ray_tracing_single_gridding`` (ray_tracing.c:484-534),
compiled in place by ``make -C oracle ref`` into oracle/_ref, evaluated for
every direction of a 24 x 12 theta/phi grid (directions and angles built as
utils_gridding.py builds them, see anacor_public_amd/gridding.py) and 48
crystal voxels of the 64^3 plain configuration, with IsExp = 0 (map entries)
and, for 8 directions, IsExp = 1.  A direction on which the reference aborts
(assert / exit) is screened in a subprocess first and recorded as such.
``gridding_interpolation.c`` needs GSL, which this image does not have; its
look-up is pinned by tests/test_oracle_gridding.py instead.

Usage:  python tests/golden/make_golden_gridding.py   (needs /root/reference)
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from anacor_public_amd import gridding as G  # noqa: E402
from anacor_public_amd import workload as W  # noqa: E402
from oracle.oracle import Reference, build, label_table  # noqa: E402

N_THETA, N_PHI, N_VOX = 24, 12, 48


def case():
    wl = W.small_config(R=1)
    coords = wl.coords[np.linspace(0, len(wl.coords), N_VOX, endpoint=False).astype(int)]
    dirs, angles = G.grid_directions(N_THETA, N_PHI)
    return wl, coords, dirs, angles


def values(d_lo: int, d_hi: int, is_exp: int):
    wl, coords, dirs, angles = case()
    ref = Reference()
    table, keep = label_table(np.ascontiguousarray(wl.volume).view(np.int8))
    out = np.empty((d_hi - d_lo, len(coords)))
    for i, d in enumerate(range(d_lo, d_hi)):
        for k, c in enumerate(coords):
            out[i, k] = ref.single_gridding(wl, c, dirs[d], angles[d, 0], angles[d, 1],
                                            is_exp=is_exp, table=table)
    del keep
    return out


def screen(d: int) -> bool:
    p = subprocess.run([sys.executable, __file__, "--screen", str(d)], capture_output=True,
                       text=True)
    return p.returncode == 0


def main() -> None:
    build(with_reference=True)
    wl, coords, dirs, angles = case()
    D = len(dirs)
    ok = np.array([screen(d) for d in range(D)])
    print(f"{int((~ok).sum())} of {D} directions abort in the reference")
    vals = np.full((D, len(coords)), np.nan)
    for d in np.flatnonzero(ok):
        vals[d] = values(d, d + 1, 0)[0]
    exp_dirs = np.flatnonzero(ok)[:: max(1, int(ok.sum()) // 8)][:8]
    vals_exp = np.stack([values(d, d + 1, 1)[0] for d in exp_dirs])
    np.savez_compressed(os.path.join(HERE, "golden_gridding.npz"), volume=wl.volume,
                        coords=coords, voxel_size=wl.voxel_size, coefficients=wl.coefficients,
                        full_iteration=wl.full_iteration, n_theta=N_THETA, n_phi=N_PHI,
                        dirs=dirs, angles=angles, ok=ok, values=vals, exp_dirs=exp_dirs,
                        values_exp=vals_exp)


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--screen":
        values(int(sys.argv[2]), int(sys.argv[2]) + 1, 0)
        sys.exit(0)
    main()
