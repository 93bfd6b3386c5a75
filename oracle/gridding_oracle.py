"""CPU restatement of the gridding look-up (TEST INFRASTRUCTURE ONLY).

The reference's nearest_neighbor_interpolate_overall (src/gridding_interpolation.c:
98-161) with its two methods:

* nearest: ``nearest_neighbor_custom`` (:16-32) -- index = (int)((x - min) / step)
  with step = (max - min) / (n - 1), clamped to the grid;
* bilinear: GSL's ``gsl_interp2d_bilinear`` as ``interpolate`` (:53-78) calls
  it.  GSL is not installed here, so this follows GSL 2.x's published
  algorithm (interpolation/bilinear.c and the binary search of
  gsl_interp_bsearch): x_array[i] <= x < x_array[i+1] by bisection, then
  (1-t)(1-u) z00 + t(1-u) z10 + (1-t)u z01 + tu z11 with t, u the fractions
  of the cell, z indexed [j * nx + i].

The map is the reference's padded layout gridding_data[N][ny][nx].  Pinned
by tests/test_oracle_gridding.py: maps that are linear in (theta, phi) (which
bilinear interpolation reproduces exactly inside the grid), grid nodes
(nearest and bilinear return the node), hand-computed cell indices, and the
padding's wrap-around.  The sum over voxels is sequential in voxel order as
in the reference.  Never imported by the product path.
"""
from __future__ import annotations

import numpy as np


def nearest_index(x, lo, hi, n):
    step = (hi - lo) / (n - 1)
    i = int((x - lo) / step)        # C truncation toward zero
    return min(max(i, 0), n - 1)


def bsearch(xs, x):
    lo, hi = 0, len(xs) - 1
    while hi > lo + 1:
        i = (hi + lo) // 2
        if xs[i] > x:
            hi = i
        else:
            lo = i
    return lo


def bilinear(grid, xs, ys, x, y):
    """GSL bilinear interp2d at (x, y); grid [..., ny, nx] (vectorised over
    leading axes); NaN outside the grid (GSL reports a domain error)."""
    if not (xs[0] <= x <= xs[-1] and ys[0] <= y <= ys[-1]):
        return np.full(grid.shape[:-2], np.nan)
    xi, yi = bsearch(xs, x), bsearch(ys, y)
    xmin, xmax, ymin, ymax = xs[xi], xs[xi + 1], ys[yi], ys[yi + 1]
    t = (x - xmin) / (xmax - xmin)
    u = (y - ymin) / (ymax - ymin)
    z00 = grid[..., yi, xi]
    z10 = grid[..., yi, xi + 1]
    z01 = grid[..., yi + 1, xi]
    z11 = grid[..., yi + 1, xi + 1]
    return (1. - t) * (1. - u) * z00 + t * (1. - u) * z10 + (1. - t) * u * z01 + t * u * z11


def lookup(gridding_data, theta_list, phi_list, angles, method: int):
    """Per reflection the mean over voxels of exp(-(m1 + m2));
    gridding_data [N][ny][nx], angles [R][4] = (theta_1, phi_1, theta, phi)."""
    g = np.asarray(gridding_data, dtype=np.float64)
    N, ny, nx = g.shape
    tl = np.asarray(theta_list, dtype=np.float64)
    pl = np.asarray(phi_list, dtype=np.float64)
    tmin, tmax, pmin, pmax = tl.min(), tl.max(), pl.min(), pl.max()
    out = np.empty(len(angles))
    for r, (t1, p1, t2, p2) in enumerate(np.asarray(angles, dtype=np.float64)):
        if method == 1:
            m1 = g[:, nearest_index(p1, pmin, pmax, ny), nearest_index(t1, tmin, tmax, nx)]
            m2 = g[:, nearest_index(p2, pmin, pmax, ny), nearest_index(t2, tmin, tmax, nx)]
        else:
            m1 = bilinear(g, tl, pl, t1, p1)
            m2 = bilinear(g, tl, pl, t2, p2)
        s = 0.0
        for v in np.exp(-(m1 + m2)):
            s += v
        out[r] = s / N
    return out
