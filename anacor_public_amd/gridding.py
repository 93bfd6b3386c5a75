"""Gridding (absorption-map) path on the GPU.

The reference's approximate method (utils/utils_gridding.py,
src/gridding_interpolation.c): for every direction of a theta/phi grid and
every crystal voxel, the sum of mu*L of one ray from the voxel along that
direction (``ray_tracing_single_gridding``, IsExp = 0); then, per reflection,
the mean over voxels of exp(-(m1 + m2)) with m1, m2 looked up at the incident
and diffracted directions (truncated nearest-neighbour index or bilinear).

* ``grid_directions`` builds the grid's direction vectors and angles with the
  same NumPy operations as utils_gridding.py:103-112, 254-259 (one scalar
  evaluation per direction, so the angles carry the same bits);
* ``padded_lists`` builds the padded theta/phi lists of utils_gridding.py:
  406-445 (the map extended by 1/12 of its extent at each end, wrapped);
* ``Gridding`` builds the map on the GPU in one launch (the exact tracer rules,
  kept resident in HBM) and corrects reflections by look-up.

There is no CPU path.
"""
from __future__ import annotations

import ctypes as ct
from typing import Tuple

import numpy as np

from . import _lib

_vp = ct.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- the grid

def thetaphi_2_myframe(theta, phi):
    """utils/utils_rt.py:321-325."""
    z = np.cos(theta) * np.sin(phi)
    x = -np.cos(theta) * np.cos(phi)
    y = -np.sin(theta)
    return np.array([z, y, x])


_MYFRAME_2_DIALS = np.linalg.inv(np.array([[1, 0, 0], [0, 0, 1], [0, 1, 0]]))


def myframe_2_dials(vector):
    """utils/utils_rt.py:483-503 (the inverse of the z/y swap)."""
    return _MYFRAME_2_DIALS.dot(vector)


def dials_2_thetaphi(rotated_s1, L1=False):
    """utils/utils_rt.py:327-351, evaluated with the same NumPy scalar
    operations (the C version is dials_2_thetaphi_22)."""
    if L1 is True:
        rotated_s1 = -rotated_s1
    if rotated_s1[1] == 0:
        theta = np.arctan(-rotated_s1[2] / (-np.sqrt(rotated_s1[0] ** 2 + rotated_s1[1] ** 2)
                                            + 0.001))
        phi = np.arctan(-rotated_s1[0] / (rotated_s1[1] + 0.001))
    else:
        if rotated_s1[1] < 0:
            theta = np.arctan(-rotated_s1[2] / np.sqrt(rotated_s1[0] ** 2 + rotated_s1[1] ** 2))
            phi = np.arctan(-rotated_s1[0] / (rotated_s1[1]))
        else:
            if rotated_s1[2] < 0:
                theta = np.pi - np.arctan(
                    -rotated_s1[2] / np.sqrt(rotated_s1[0] ** 2 + rotated_s1[1] ** 2))
            else:
                theta = -np.pi - np.arctan(
                    -rotated_s1[2] / np.sqrt(rotated_s1[0] ** 2 + rotated_s1[1] ** 2))
            phi = -np.arctan(-rotated_s1[0] / (-rotated_s1[1]))
    return theta, phi


def grid_directions(n_theta: int, n_phi: int) -> Tuple[np.ndarray, np.ndarray]:
    """(rotated_s1 [D][3], theta_phi [D][2]) of the n_theta x n_phi grid,
    phi-major (utils_gridding.py:254-259 for the grid, :103-112 and :161-163
    for the vectors and their angles)."""
    theta_list = np.linspace(-180, 180, n_theta, endpoint=False)
    phi_list = np.linspace(-90, 90, n_phi, endpoint=False)
    theta_grid, phi_grid = np.meshgrid(theta_list, phi_list)
    data = np.stack((theta_grid.ravel(), phi_grid.ravel()), axis=-1)
    data = data / 180 * np.pi
    vecs = np.empty((len(data), 3))
    angles = np.empty((len(data), 2))
    for i, (theta, phi) in enumerate(data):
        v = myframe_2_dials(thetaphi_2_myframe(theta, phi))
        vecs[i] = v
        angles[i] = dials_2_thetaphi(v)
    return vecs, angles


def padding(n_theta: int, n_phi: int) -> Tuple[int, int]:
    """Columns and rows added at each end (utils_gridding.py:406-407)."""
    return int(np.round((n_theta // 6) / 2)), int(np.round(n_phi // 6 / 2))


def padded_lists(n_theta: int, n_phi: int) -> Tuple[np.ndarray, np.ndarray]:
    """theta_list and phi_list of the padded map (utils_gridding.py:424-445)."""
    te, pe = padding(n_theta, n_phi)
    len_theta = int(n_theta + te * 2)
    len_phi = int(n_phi + pe * 2)
    theta_list = (np.linspace(-np.round(180 * 7 / 6), np.round(180 * 7 / 6), len_theta,
                              endpoint=False) / 180 * np.pi)
    phi_list = (np.linspace(-np.round(90 * 7 / 6), np.round(90 * 7 / 6), len_phi,
                            endpoint=False) / 180 * np.pi)
    return theta_list, phi_list


def pad_map(abs_map: np.ndarray, n_theta: int, n_phi: int) -> np.ndarray:
    """The reference's padded map, [N][ny][nx] float64 (utils_gridding.py:
    390-424), from a [D][N] map; for the drop-in look-up."""
    g = np.asarray(abs_map, dtype=np.float64).reshape((n_phi, n_theta, -1))
    te, pe = padding(n_theta, n_phi)
    if te:
        g = np.concatenate((g[:, -te:, :], g, g[:, :te, :]), axis=1)
    if pe:
        g = np.concatenate((g[-pe:, :, :], g, g[:pe, :, :]), axis=0)
    return np.ascontiguousarray(np.transpose(g, (2, 0, 1)))


# ---------------------------------------------------------------- the path

class Gridding:
    """A resident absorption map of one sample's voxels (the Tracer's volume
    and voxel list) for an n_theta x n_phi grid."""

    NEAREST = 1
    BILINEAR = 2

    def __init__(self, tracer, n_theta: int = 360, n_phi: int = 180, is_exp: bool = False):
        self._lib = _lib.load()
        self.tracer = tracer
        self.n_theta, self.n_phi = int(n_theta), int(n_phi)
        self.directions, self.angles = grid_directions(self.n_theta, self.n_phi)
        self.theta_list, self.phi_list = padded_lists(self.n_theta, self.n_phi)
        h = self._lib.anacor_grid_create(tracer.handle, self.n_theta, self.n_phi,
                                         _p(_f64(self.directions)), _p(_f64(self.angles)),
                                         1 if is_exp else 0)
        if not h:
            code, msg = _lib.last_error(self._lib)
            raise _lib.AnacorError(code, msg)
        self._h = ct.c_void_p(h)

    @property
    def handle(self):
        if not getattr(self, "_h", None):
            raise ValueError("Gridding is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.anacor_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def map(self) -> np.ndarray:
        """float32 [D][N]: the reference's absorption_map layout."""
        out = np.empty((self.n_theta * self.n_phi, self.tracer.n_voxels), dtype=np.float32)
        _lib.check(self._lib.anacor_grid_map(self.handle, _p(out)), self._lib)
        return out

    def run(self, s1, omega, raw_xray, omega_axis, F, method: int = NEAREST,
            strict: bool = False) -> np.ndarray:
        s1 = _f64(s1).reshape(-1, 3)
        om = _f64(omega).reshape(-1)
        out = np.empty(len(om), dtype=np.float64)
        rc = self._lib.anacor_grid_run(self.handle, _p(s1), _p(om), len(om), _p(_f64(raw_xray)),
                                       _p(_f64(omega_axis)), _p(_f64(F)), _p(self.theta_list),
                                       len(self.theta_list), _p(self.phi_list),
                                       len(self.phi_list), int(method), _p(out))
        _lib.check(rc, self._lib, allow=(_lib.ANACOR_OK,) if strict else
                   (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY))
        return out

    def stats(self) -> dict:
        b, l = ct.c_double(0.0), ct.c_double(0.0)
        self._lib.anacor_grid_stats(self.handle, ct.byref(b), ct.byref(l))
        return {"build_ms": b.value, "lookup_ms": l.value}


def nearest_neighbor_interpolate_overall(coord_list, s1, omega, raw_xray, omega_axis, F,
                                         voxel_size, coefficients, theta_list, phi_list,
                                         gridding_data, method: int = 1,
                                         num_workers: int = 0) -> np.ndarray:
    """The drop-in C entry with the argument layout of utils_gridding.py:
    627-641 (gridding_data: the padded [N][ny][nx] map)."""
    lib = _lib.load()
    coords = np.ascontiguousarray(coord_list, dtype=np.int64)
    s1 = _f64(s1).reshape(-1, 3)
    om = _f64(omega).reshape(-1)
    tl, pl = _f64(theta_list), _f64(phi_list)
    grid = _f64(gridding_data)
    p = lib.nearest_neighbor_interpolate_overall(
        0, len(om), _p(coords), len(coords), _p(s1), _p(om), _p(_f64(raw_xray)),
        _p(_f64(omega_axis)), _p(_f64(F).reshape(-1)), len(om), _p(_f64(voxel_size)),
        _p(_f64(coefficients)), int(num_workers), 1, _p(tl), _p(pl), _p(grid), len(tl), len(pl),
        float(tl.min()), float(tl.max()), float(pl.min()), float(pl.max()), int(method))
    out = np.ctypeslib.as_array(p, shape=(len(om),)).copy()
    lib.free(p)
    return out
