"""Host-side interface of the MI355X ray tracer, mirroring the reference's.

* ``ray_tracing_overall`` / ``ray_tracing_gpu_overall`` / ``ray_tracing_single``
  call the C-ABI entry points with the argument layout the reference's
  ``worker_function`` uses (utils/utils_mp.py:59-79, :128-148, :178-204) and
  return numpy arrays (the library's malloc'd result is freed with libc
  ``free`` as utils_mp.py:212 does).
* ``python_2_c_3d`` and ``kp_rotation`` mirror utils/utils_os.py:26-68.
* ``Tracer`` keeps the label volume and the voxel list resident in HBM and
  traces reflection batches against them (the library's extended interface).

There is no CPU fallback: every call goes through libanacor_mi355x.so and
raises if it is missing or no HIP device is visible.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional

import numpy as np

from . import _lib

_vp = ct.c_void_p


def _p(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(-1) if n is None else a.reshape(n)


def kp_rotation(axis, theta):
    """Rodrigues rotation matrix (utils/utils_os.py:26-43)."""
    x, y, z = axis
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c + (x ** 2) * (1 - c), x * y * (1 - c) - z * s, y * s + x * z * (1 - c)],
                     [z * s + x * y * (1 - c), c + (y ** 2) * (1 - c), -x * s + y * z * (1 - c)],
                     [-y * s + x * z * (1 - c), x * s + y * z * (1 - c), c + (z ** 2) * (1 - c)]])


class LabelTable:
    """The int8*** row-pointer table of ``python_2_c_3d`` (utils/utils_os.py:45-68),
    built with numpy; keeps the volume and the pointer arrays alive."""

    def __init__(self, label_list: np.ndarray):
        vol = np.ascontiguousarray(label_list)
        if vol.dtype not in (np.int8, np.uint8) or vol.ndim != 3:
            raise ValueError("label_list must be a 3-D int8/uint8 array")
        Z, Y, X = vol.shape
        self.volume = vol
        self.rows = (vol.ctypes.data + np.arange(Z * Y, dtype=np.uint64) * np.uint64(X)).astype(np.uint64)
        self.planes = (self.rows.ctypes.data + np.arange(Z, dtype=np.uint64) * np.uint64(Y * 8)).astype(np.uint64)
        self.ptr = ct.c_void_p(self.planes.ctypes.data)


def python_2_c_3d(label_list: np.ndarray) -> LabelTable:
    return LabelTable(label_list)


def ray_tracing_overall(low, up, coord_list, len_coord_list, scattering_vector_list, omega_list,
                        raw_xray, omega_axis, kp_rotation_matrix, len_result, voxel_size,
                        coefficients, label_list, shape, full_iteration=0, store_paths=0,
                        num_workers=8, IsExp=1) -> np.ndarray:
    """Drop-in for the reference's FP64 entry (ray_tracing.c:365-425).
    ``label_list`` may be a LabelTable or a 3-D array.  Returns float64[len_result];
    reflections the reference would abort on are NaN (see ``last_error``)."""
    lib = _lib.load()
    table = label_list if hasattr(label_list, "ptr") else LabelTable(label_list)
    coords = np.ascontiguousarray(coord_list, dtype=np.int64)
    s1 = _f64(scattering_vector_list)
    om = _f64(omega_list)
    p = lib.ray_tracing_overall(int(low), int(up), _p(coords), int(len_coord_list), _p(s1), _p(om),
                                _p(_f64(raw_xray)), _p(_f64(omega_axis)),
                                _p(_f64(kp_rotation_matrix)), int(len_result), _p(_f64(voxel_size)),
                                _p(_f64(coefficients)), table.ptr,
                                _p(np.ascontiguousarray(shape, dtype=np.int64)), int(full_iteration),
                                int(store_paths), int(num_workers), int(IsExp))
    if not p:
        raise MemoryError("ray_tracing_overall returned NULL")
    out = np.ctypeslib.as_array(p, shape=(int(len_result),)).copy()
    lib.free(ct.cast(p, _vp))
    code, msg = _lib.last_error(lib)
    if code not in (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY):
        raise _lib.AnacorError(code, msg)
    return out


def ray_tracing_gpu_overall(low, up, coord_list, len_coord_list, scattering_vector_list,
                            omega_list, raw_xray, omega_axis, kp_rotation_matrix, len_result,
                            voxel_size, coefficients, label_list, shape, full_iteration=0,
                            store_paths=0, gpumethod=1) -> np.ndarray:
    """Drop-in for the reference's FP32 GPU entry (ray_tracing_gpu.c:33-62):
    float32 inputs, int32 coordinates and shape.  Returns float32[len_result]."""
    lib = _lib.load()
    table = label_list if hasattr(label_list, "ptr") else LabelTable(label_list)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(-1)  # noqa: E731
    coords = np.ascontiguousarray(coord_list, dtype=np.int32)
    p = lib.ray_tracing_gpu_overall(int(low), int(up), _p(coords), int(len_coord_list),
                                    _p(f32(scattering_vector_list)), _p(f32(omega_list)),
                                    _p(f32(raw_xray)), _p(f32(omega_axis)),
                                    _p(f32(kp_rotation_matrix)), int(len_result),
                                    _p(f32(voxel_size)), _p(f32(coefficients)), table.ptr,
                                    _p(table.volume), _p(np.ascontiguousarray(shape, dtype=np.int32)),
                                    int(full_iteration), int(store_paths), int(gpumethod))
    if not p:
        raise MemoryError("ray_tracing_gpu_overall returned NULL")
    out = np.ctypeslib.as_array(p, shape=(int(len_result),)).copy()
    lib.free(ct.cast(p, _vp))
    code, msg = _lib.last_error(lib)
    if code not in (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY):
        raise _lib.AnacorError(code, msg)
    return out


def ray_tracing_single(coord_list, len_coord_list, rotated_s1, xray, voxel_size, coefficients,
                       label_list, shape, full_iteration=0, store_paths=0, IsExp=1) -> float:
    """Drop-in for ray_tracing_single (ray_tracing.c:166-292)."""
    lib = _lib.load()
    table = label_list if hasattr(label_list, "ptr") else LabelTable(label_list)
    coords = np.ascontiguousarray(coord_list, dtype=np.int64)
    v = lib.ray_tracing_single(_p(coords), int(len_coord_list), _p(_f64(rotated_s1)), _p(_f64(xray)),
                               _p(_f64(voxel_size)), _p(_f64(coefficients)), table.ptr,
                               _p(np.ascontiguousarray(shape, dtype=np.int64)), int(full_iteration),
                               int(store_paths), int(IsExp))
    code, msg = _lib.last_error(lib)
    if code not in (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY):
        raise _lib.AnacorError(code, msg)
    return float(v)


def ray_tracing_single_mp(coord_list, len_coord_list, rotated_s1, xray, voxel_size, coefficients,
                          label_list, shape, full_iteration=0, store_paths=0, IsExp=1,
                          num_workers=1) -> float:
    """Drop-in for ray_tracing_single_mp (ray_tracing.c:294-364)."""
    lib = _lib.load()
    table = label_list if hasattr(label_list, "ptr") else LabelTable(label_list)
    coords = np.ascontiguousarray(coord_list, dtype=np.int64)
    v = lib.ray_tracing_single_mp(_p(coords), int(len_coord_list), _p(_f64(rotated_s1)),
                                  _p(_f64(xray)), _p(_f64(voxel_size)), _p(_f64(coefficients)),
                                  table.ptr, _p(np.ascontiguousarray(shape, dtype=np.int64)),
                                  int(full_iteration), int(store_paths), int(IsExp),
                                  int(num_workers))
    code, msg = _lib.last_error(lib)
    if code not in (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY):
        raise _lib.AnacorError(code, msg)
    return float(v)


def ib_test(coord_list, len_coord_list, rotated_s1, xray, voxel_size, coefficients, label_list,
            shape, full_iteration=0, store_paths=0, num_cls=3, IsExp=1) -> float:
    """Drop-in for ib_test (ray_tracing.c:25-160): the bisection traversal of
    one reflection given its rotated vectors.  IsExp is accepted and ignored,
    as by the reference (its binding passes it, the C function has no such
    parameter and always averages exp(-sum mu*L))."""
    lib = _lib.load()
    table = label_list if hasattr(label_list, "ptr") else LabelTable(label_list)
    coords = np.ascontiguousarray(coord_list, dtype=np.int64)
    v = lib.ib_test(_p(coords), int(len_coord_list), _p(_f64(rotated_s1)), _p(_f64(xray)),
                    _p(_f64(voxel_size)), _p(_f64(coefficients)), table.ptr,
                    _p(np.ascontiguousarray(shape, dtype=np.int64)), int(full_iteration),
                    int(store_paths), int(num_cls))
    code, msg = _lib.last_error(lib)
    if code not in (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY):
        raise _lib.AnacorError(code, msg)
    return float(v)


def num_classes(volume) -> int:
    """Number of labels in the volume minus one (the vacuum), as the
    reference's driver counts num_cls for ib_test (main_lite.py:153)."""
    vol = np.ascontiguousarray(volume).view(np.uint8).reshape(-1)
    return int(np.count_nonzero(np.bincount(vol, minlength=256))) - 1


def last_error():
    return _lib.last_error()


class Tracer:
    """Label volume and crystal-voxel list resident on one GPU.

    >>> t = Tracer(volume, coords, voxel_size, coefficients)
    >>> A = t.run(s1, omega, raw_xray, omega_axis, F)      # float64[R]
    """

    def __init__(self, volume: np.ndarray, coords: np.ndarray, voxel_size=(1.0, 1.0, 1.0),
                 coefficients=(0.0, 0.0, 0.0, 0.0), full_iteration: int = 0, device: int = 0):
        self._lib = _lib.load()
        vol = np.ascontiguousarray(volume)
        if vol.ndim != 3:
            raise ValueError("volume must be 3-D [z][y][x]")
        vol = vol.view(np.uint8) if vol.dtype in (np.int8, np.uint8) else vol.astype(np.uint8)
        self.shape = np.array(vol.shape, dtype=np.int64)
        self.coords = np.ascontiguousarray(coords, dtype=np.int64).reshape(-1, 3)
        self.n_voxels = len(self.coords)
        h = self._lib.anacor_ctx_create(_p(vol), _p(self.shape), _p(self.coords), self.n_voxels,
                                        _p(_f64(voxel_size)), _p(_f64(coefficients)),
                                        int(full_iteration), int(device))
        if not h:
            code, msg = _lib.last_error(self._lib)
            raise _lib.AnacorError(code, msg)
        self._h = ct.c_void_p(h)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.anacor_ctx_destroy(self._h)
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

    def set_tile(self, voxels_per_tile: int):
        self._lib.anacor_ctx_set_tile(self.handle, int(voxels_per_tile))

    def set_order(self, refl_group: int = 0, direction_major: int = -1):
        """Launch order of the trace kernel: groups of refl_group direction-sorted
        reflections run tile by tile (0 automatic, 1 none); the sort key is led
        by the incident (0) or the diffracted (1) ray (-1 automatic).  Speed only."""
        _lib.check(self._lib.anacor_ctx_set_order(self.handle, int(refl_group),
                                                  int(direction_major)), self._lib)

    def work_shape(self) -> dict:
        """The work shape in use: voxel block, tile setting (0 automatic),
        reflections per launch group, leading ray of the direction sort."""
        blk = (ct.c_int * 3)()
        tile, group, major = ct.c_int(), ct.c_int(), ct.c_int()
        _lib.check(self._lib.anacor_ctx_get_work_shape(self.handle, blk, ct.byref(tile),
                                                       ct.byref(group), ct.byref(major)),
                   self._lib)
        return {"block": list(blk), "tile": tile.value, "group": group.value,
                "direction_major": major.value}

    def set_layout(self, layout: int):
        """Volume layout of the trace kernel (_lib.LAYOUT_*): raw uint8, or 8x8x8
        bricks of 2-bit / 4-bit labels.  Results are bit-identical in every
        layout; bricks pay off when the volume outgrows the caches."""
        _lib.check(self._lib.anacor_ctx_set_layout(self.handle, int(layout)), self._lib)

    @property
    def layout(self) -> int:
        """Layout used by the most recent trace (_lib.LAYOUT_RAW, _BRICK2, _BRICK4)."""
        return int(self._lib.anacor_ctx_get_layout(self.handle))

    def set_voxel_block(self, bz: int, by: int, bx: int):
        """Group crystal voxels into waves by (bz, by, bx) voxel blocks
        ((0, 0, 0): Morton order)."""
        _lib.check(self._lib.anacor_ctx_set_voxel_block(self.handle, int(bz), int(by), int(bx)),
                   self._lib)

    def run(self, s1, omega, raw_xray, omega_axis, F, f64_mean: bool = False, is_exp: bool = True,
            strict: bool = False) -> np.ndarray:
        s1 = _f64(s1).reshape(-1, 3)
        om = _f64(omega)
        R = len(om)
        out = np.empty(R, dtype=np.float64)
        flags = (_lib.FLAG_F64_MEAN if f64_mean else 0) | (0 if is_exp else _lib.FLAG_NO_EXP)
        rc = self._lib.anacor_ctx_run(self.handle, _p(s1), _p(om), R, _p(_f64(raw_xray)),
                                      _p(_f64(omega_axis)), _p(_f64(F)), flags, _p(out))
        _lib.check(rc, self._lib, allow=(_lib.ANACOR_OK,) if strict else
                   (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY))
        return out

    def paths(self, s1, omega, raw_xray, omega_axis, F) -> np.ndarray:
        """Per (reflection, voxel) path records, shape [R, N] of PATH_DTYPE."""
        s1 = _f64(s1).reshape(-1, 3)
        om = _f64(omega)
        R = len(om)
        out = np.zeros(R * self.n_voxels, dtype=_lib.PATH_DTYPE)
        rc = self._lib.anacor_ctx_paths(self.handle, _p(s1), _p(om), R, _p(_f64(raw_xray)),
                                        _p(_f64(omega_axis)), _p(_f64(F)), _p(out))
        _lib.check(rc, self._lib)
        return out.reshape(R, self.n_voxels)

    def bisection(self, s1, omega, raw_xray, omega_axis, F, num_cls: int = 3,
                  f64_mean: bool = False, is_exp: bool = True,
                  strict: bool = False) -> np.ndarray:
        """The reference's iterative-bisection traversal (ib_test after the
        rotation of ray_tracing_overall) for every reflection in one launch."""
        s1 = _f64(s1).reshape(-1, 3)
        om = _f64(omega)
        R = len(om)
        out = np.empty(R, dtype=np.float64)
        flags = (_lib.FLAG_F64_MEAN if f64_mean else 0) | (0 if is_exp else _lib.FLAG_NO_EXP)
        rc = self._lib.anacor_ctx_bisection(self.handle, _p(s1), _p(om), R, _p(_f64(raw_xray)),
                                            _p(_f64(omega_axis)), _p(_f64(F)), int(num_cls),
                                            flags, _p(out))
        _lib.check(rc, self._lib, allow=(_lib.ANACOR_OK,) if strict else
                   (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY))
        return out

    def bisection_rotated(self, rotated_s1, xray, num_cls: int = 3, f64_mean: bool = False,
                          is_exp: bool = True, lengths: bool = False):
        """ib_test for R reflections given as rotated vectors ([R][3] each).
        With lengths=True also returns [R, N, 2, 4] per-voxel, per-ray
        [li, lo, cr, bu] (voxels in the order given to the constructor)."""
        s1 = _f64(rotated_s1).reshape(-1, 3)
        xr = _f64(xray).reshape(-1, 3)
        if len(s1) != len(xr):
            raise ValueError("rotated_s1 and xray differ in length")
        R = len(s1)
        out = np.empty(R, dtype=np.float64)
        L = np.empty((R, self.n_voxels, 2, 4), dtype=np.float64) if lengths else None
        flags = (_lib.FLAG_F64_MEAN if f64_mean else 0) | (0 if is_exp else _lib.FLAG_NO_EXP)
        rc = self._lib.anacor_ctx_bisection_rotated(self.handle, _p(s1), _p(xr), R, int(num_cls),
                                                    flags, _p(out), _p(L) if lengths else None)
        _lib.check(rc, self._lib, allow=(_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY))
        return (out, L) if lengths else out

    def stats(self) -> dict:
        st = _lib.Stats()
        self._lib.anacor_ctx_stats(self.handle, ct.byref(st))
        return dict(kernel_ms=st.kernel_ms, voxel_steps=int(st.voxel_steps),
                    workgroups=int(st.workgroups), n_voxels=int(st.n_voxels),
                    host_ms=st.host_ms, outside_reads=int(st.outside_reads))

    @classmethod
    def _borrow(cls, handle: int, n_voxels: int, shape) -> "Tracer":
        """A non-owning Tracer over a handle owned by a Batch."""
        t = cls.__new__(cls)
        t._lib = _lib.load()
        t._h = None
        t._borrowed = ct.c_void_p(handle)
        t.n_voxels = n_voxels
        t.shape = np.array(shape, dtype=np.int64)
        return t

    @property
    def handle(self):
        return self._h if self._h is not None else getattr(self, "_borrowed", None)

    @classmethod
    def for_workload(cls, wl, device: int = 0) -> "Tracer":
        return cls(wl.volume, wl.coords, wl.voxel_size, wl.coefficients, wl.full_iteration, device)

    def run_workload(self, wl, **kw) -> np.ndarray:
        return self.run(wl.s1, wl.omega, wl.raw_xray, wl.omega_axis, wl.F, **kw)


def reflection_geometry(s1, omega, raw_xray, omega_axis, F, max_dim: int):
    """Host-side (theta, phi) per ray and the per-face delta table, as the
    library computes them (diagnostics for the parity tests)."""
    lib = _lib.load()
    s1 = _f64(s1).reshape(-1, 3)
    om = _f64(omega)
    R = len(om)
    tp = np.zeros((R, 4))
    dt = np.zeros((R, 2, 6, 3))
    _lib.check(lib.anacor_reflection_geometry(_p(s1), _p(om), R, _p(_f64(raw_xray)),
                                              _p(_f64(omega_axis)), _p(_f64(F)), int(max_dim),
                                              _p(tp), _p(dt)), lib)
    return tp, dt


class Batch:
    """Several independent samples traced together (``anacor_batch_*``): per
    sample its host setup, rule upload and trace launch, each setup
    overlapping the trace before it; launched from the host or, with
    ``set_graph(True)``, as one hipGraph.  ``samples`` are Workload-like
    objects (volume, coords, voxel_size, coefficients); all share
    full_iteration."""

    def __init__(self, samples, full_iteration: int = 0, device: int = 0):
        self._lib = _lib.load()
        self._keep = []
        arr = (_lib.Sample * len(samples))()
        for i, wl in enumerate(samples):
            vol = np.ascontiguousarray(wl.volume).view(np.uint8)
            co = np.ascontiguousarray(wl.coords, dtype=np.int64).reshape(-1, 3)
            self._keep += [vol, co]
            arr[i].volume = vol.ctypes.data
            arr[i].shape[:] = list(vol.shape)
            arr[i].coords = co.ctypes.data
            arr[i].n_coords = len(co)
            arr[i].voxel_size[:] = [float(v) for v in wl.voxel_size]
            arr[i].coefficients[:] = [float(v) for v in wl.coefficients]
        h = self._lib.anacor_batch_create(arr, len(samples), int(full_iteration), int(device))
        if not h:
            code, msg = _lib.last_error(self._lib)
            raise _lib.AnacorError(code, msg)
        self._h = ct.c_void_p(h)
        self.n = len(samples)
        self._sizes = [(len(np.asarray(w.coords).reshape(-1, 3)), np.asarray(w.volume).shape)
                       for w in samples]
        self._keep = []   # the library copied volumes and voxel lists to the GPU

    def sample(self, i: int) -> Tracer:
        h = self._lib.anacor_batch_sample(self._h, int(i))
        if not h:
            raise IndexError(i)
        return Tracer._borrow(h, self._sizes[i][0], self._sizes[i][1])

    def run(self, s1_list, omega_list, raw_xray, omega_axis, F, f64_mean: bool = False,
            is_exp: bool = True, strict: bool = False):
        n = self.n
        s1s = [_f64(a).reshape(-1, 3) for a in s1_list]
        oms = [_f64(a) for a in omega_list]
        outs = [np.empty(len(o), dtype=np.float64) for o in oms]
        p_s1 = (ct.c_void_p * n)(*[a.ctypes.data for a in s1s])
        p_om = (ct.c_void_p * n)(*[a.ctypes.data for a in oms])
        p_out = (ct.c_void_p * n)(*[a.ctypes.data for a in outs])
        Rs = np.array([len(o) for o in oms], dtype=np.int64)
        flags = (_lib.FLAG_F64_MEAN if f64_mean else 0) | (0 if is_exp else _lib.FLAG_NO_EXP)
        rc = self._lib.anacor_batch_run(self._h, ct.cast(p_s1, _vp), ct.cast(p_om, _vp), _p(Rs),
                                        _p(_f64(raw_xray)), _p(_f64(omega_axis)), _p(_f64(F)),
                                        flags, ct.cast(p_out, _vp))
        _lib.check(rc, self._lib, allow=(_lib.ANACOR_OK,) if strict else
                   (_lib.ANACOR_OK, _lib.ANACOR_ERR_GEOMETRY))
        return outs

    def set_graph(self, use_graph: bool) -> None:
        """Run as one hipGraph (True) or launch from the host (False, default)."""
        _lib.check(self._lib.anacor_batch_set_graph(self._h, int(bool(use_graph))), self._lib)

    @property
    def last_graph(self) -> bool:
        """Whether the most recent run launched the batch graph."""
        return bool(self._lib.anacor_batch_last_graph(self._h))

    def stats(self) -> dict:
        st = _lib.Stats()
        self._lib.anacor_batch_stats(self._h, ct.byref(st))
        return dict(kernel_ms=st.kernel_ms, voxel_steps=int(st.voxel_steps),
                    workgroups=int(st.workgroups), n_voxels=int(st.n_voxels),
                    host_ms=st.host_ms)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.anacor_batch_destroy(self._h)
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
