"""ctypes bindings for the test oracle.

TEST INFRASTRUCTURE ONLY.  ``load_oracle()`` loads the FP64 CPU restatement
(``oracle/liboracle_anacor.so``); ``load_reference()`` loads the reference's own
C sources compiled in place by ``make -C oracle ref`` (``oracle/_ref``).  Only
tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
import this module; the product library never does.
"""
from __future__ import annotations

import contextlib
import ctypes as ct
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ANACOR_ORACLE_LIB selects another build of the restatement (e.g. the
# sanitizer build of `make -C oracle sanitize`)
ORACLE_SO = os.environ.get("ANACOR_ORACLE_LIB", os.path.join(HERE, "liboracle_anacor.so"))
REF_SO = os.path.join(HERE, "_ref", "ray_tracing_cpu.so")
REF_SRC = "/root/reference/AnACor/src"


class OracleRay(ct.Structure):
    _fields_ = [("face", ct.c_int32), ("err", ct.c_int32), ("len", ct.c_int64),
                ("nseg", ct.c_int64), ("first", ct.c_int64 * 3), ("last", ct.c_int64 * 3),
                ("L", ct.c_double * 4)]


RAY_DTYPE = np.dtype([("face", np.int32), ("err", np.int32), ("len", np.int64),
                      ("nseg", np.int64), ("first", np.int64, 3), ("last", np.int64, 3),
                      ("L", np.float64, 4)])
assert RAY_DTYPE.itemsize == ct.sizeof(OracleRay)


class OracleRefl(ct.Structure):
    _fields_ = [("xray", ct.c_double * 3), ("s1", ct.c_double * 3), ("theta", ct.c_double * 2),
                ("phi", ct.c_double * 2), ("dir", (ct.c_double * 3) * 2)]


REFL_DTYPE = np.dtype([("xray", np.float64, 3), ("s1", np.float64, 3), ("theta", np.float64, 2),
                       ("phi", np.float64, 2), ("dir", np.float64, (2, 3))])


class IbDiag(ct.Structure):
    _fields_ = [("reads", ct.c_int64), ("oob", ct.c_int64), ("undefined", ct.c_int64)]


class OracleRule(ct.Structure):
    _fields_ = [("d", ct.c_double * 3), ("rounding", ct.c_int32), ("axis", ct.c_int32),
                ("to_max", ct.c_int32), ("inclusive", ct.c_int32), ("valid", ct.c_int32),
                ("vacuum_run", ct.c_int32)]


def build(with_reference: bool = True) -> None:
    """Compile the restatement and, when the reference sources are present,
    the reference build under oracle/_ref."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle_anacor.so"], check=True)
    if with_reference and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_P = np.ctypeslib.ndpointer


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


class Oracle:
    """The FP64 restatement."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing; run `make -C oracle`")
        lib = ct.CDLL(path)
        lib.oracle_trace.restype = ct.c_int64
        lib.oracle_trace.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int64,
                                     ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                     ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int,
                                     ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int64)]
        lib.oracle_single.restype = ct.c_double
        lib.oracle_single.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                      ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int,
                                      ct.c_void_p, ct.POINTER(ct.c_int64)]
        lib.oracle_cube_face.restype = ct.c_int64
        lib.oracle_cube_face.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int]
        lib.oracle_thetaphi.restype = None
        lib.oracle_thetaphi.argtypes = [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
        lib.oracle_face_rule.restype = None
        lib.oracle_face_rule.argtypes = [ct.c_double, ct.c_double, ct.c_int64, ct.POINTER(OracleRule)]
        lib.oracle_reflection_setup.restype = None
        lib.oracle_reflection_setup.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p,
                                                ct.c_void_p, ct.c_void_p, ct.c_void_p]
        lib.oracle_kp_rotation.restype = None
        lib.oracle_kp_rotation.argtypes = [ct.c_void_p, ct.c_double, ct.c_void_p]
        lib.oracle_single_ray.restype = ct.c_double
        lib.oracle_single_ray.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_double,
                                          ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                          ct.c_int, ct.c_int, ct.c_void_p]
        lib.oracle_gridding_map.restype = None
        lib.oracle_gridding_map.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                                            ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                            ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]
        lib.oracle_ib_ratios.restype = ct.c_int
        lib.oracle_ib_ratios.argtypes = [ct.c_int64, ct.c_double, ct.c_double, ct.c_void_p]
        lib.oracle_ib_test.restype = ct.c_double
        lib.oracle_ib_test.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                                       ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                       ct.c_int, ct.c_int, ct.c_void_p, ct.POINTER(IbDiag)]
        lib.oracle_ib_overall.restype = ct.c_int64
        lib.oracle_ib_overall.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                                          ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                          ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                          ct.c_int, ct.c_int, ct.c_int, ct.c_void_p,
                                          ct.POINTER(IbDiag)]
        lib.oracle_quot_rn_check.restype = ct.c_int64
        lib.oracle_quot_rn_check.argtypes = [ct.c_int, ct.c_int64, ct.c_uint64,
                                             ct.POINTER(ct.c_int64)]
        self.lib = lib

    def quot_rn_check(self, kind: int, n: int, seed: int):
        """(mismatches of two fused corrections, of one) of the GPU's quotient
        sequence against IEEE division on n seeded cases (kind 0: integer /
        double, 1: integer span / integer D)."""
        one = ct.c_int64(0)
        bad = self.lib.oracle_quot_rn_check(int(kind), int(n), int(seed), ct.byref(one))
        return int(bad), int(one.value)

    def ib_ratios(self, face, theta, phi):
        """increments (bisection.c:21-148): (rz, ry, rx) or None where the
        reference's assert fails."""
        r = np.zeros(3)
        ok = self.lib.oracle_ib_ratios(int(face), float(theta), float(phi), _ptr(r))
        return r if ok else None

    def ib_test(self, wl, coords, rotated_s1, xray, num_cls=3, is_exp=1, lengths=False):
        """ib_test (ray_tracing.c:25-160) on rotated vectors.  Returns
        (mean, diag dict[, lengths [n][2][4]])."""
        coords = np.ascontiguousarray(coords, dtype=np.int64)
        vol = np.ascontiguousarray(wl.volume).view(np.uint8)
        shape = np.array(vol.shape, dtype=np.int64)
        L = np.zeros((len(coords), 2, 4)) if lengths else None
        d = IbDiag()
        m = self.lib.oracle_ib_test(_ptr(coords), len(coords),
                                    _ptr(np.ascontiguousarray(rotated_s1, dtype=np.float64)),
                                    _ptr(np.ascontiguousarray(xray, dtype=np.float64)),
                                    _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
                                    _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)),
                                    _ptr(vol), _ptr(shape), int(num_cls), int(is_exp),
                                    _ptr(L) if lengths else None, ct.byref(d))
        diag = dict(reads=d.reads, oob=d.oob, undefined=d.undefined)
        return (m, diag, L) if lengths else (m, diag)

    def ib_overall(self, wl, num_cls=3, is_exp=1, n_threads=0, R=None):
        """Rotation then ib_test for every reflection of a Workload."""
        s1 = np.ascontiguousarray(wl.s1, dtype=np.float64)
        om = np.ascontiguousarray(wl.omega, dtype=np.float64)
        if R is not None:
            s1, om = s1[:R], om[:R]
        coords = np.ascontiguousarray(wl.coords, dtype=np.int64)
        vol = np.ascontiguousarray(wl.volume).view(np.uint8)
        shape = np.array(vol.shape, dtype=np.int64)
        out = np.empty(len(om))
        d = IbDiag()
        failed = self.lib.oracle_ib_overall(
            _ptr(coords), len(coords), _ptr(s1), _ptr(om), len(om),
            _ptr(np.ascontiguousarray(wl.raw_xray, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.omega_axis, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.F, dtype=np.float64).reshape(-1)),
            _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)), _ptr(vol), _ptr(shape),
            int(num_cls), int(is_exp), int(n_threads or os.cpu_count() or 1), _ptr(out),
            ct.byref(d))
        return dict(result=out, failed=int(failed), reads=d.reads, oob=d.oob,
                    undefined=d.undefined)

    def trace(self, wl, n_threads: int = 0, rays: bool = False, R: int | None = None):
        """A_h per reflection (and optionally per-ray records) for a Workload."""
        s1 = np.ascontiguousarray(wl.s1, dtype=np.float64)
        om = np.ascontiguousarray(wl.omega, dtype=np.float64)
        if R is not None:
            s1, om = s1[:R], om[:R]
        R_ = len(om)
        coords = np.ascontiguousarray(wl.coords, dtype=np.int64)
        N = len(coords)
        vol = np.ascontiguousarray(wl.volume).view(np.uint8)
        shape = np.array(vol.shape, dtype=np.int64)
        out = np.empty(R_, dtype=np.float64)
        rr = np.zeros(R_ * N * 2, dtype=RAY_DTYPE) if rays else None
        steps = ct.c_int64(0)
        nt = n_threads or os.cpu_count() or 1
        failed = self.lib.oracle_trace(
            _ptr(coords), N, _ptr(s1), _ptr(om), R_,
            _ptr(np.ascontiguousarray(wl.raw_xray, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.omega_axis, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.F, dtype=np.float64).reshape(-1)),
            _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)),
            _ptr(vol), _ptr(shape), int(wl.full_iteration), 1, nt,
            _ptr(out), _ptr(rr) if rays else None, ct.byref(steps))
        res = dict(result=out, failed=int(failed), steps=int(steps.value))
        if rays:
            res["rays"] = rr.reshape(R_, N, 2)
        return res

    def single(self, coords, rotated_s1, xray, voxel_size, coefficients, volume,
               full_iteration=0, is_exp=1, rays=False):
        coords = np.ascontiguousarray(coords, dtype=np.int64)
        vol = np.ascontiguousarray(volume).view(np.uint8)
        shape = np.array(vol.shape, dtype=np.int64)
        rr = np.zeros(len(coords) * 2, dtype=RAY_DTYPE) if rays else None
        steps = ct.c_int64(0)
        m = self.lib.oracle_single(_ptr(coords), len(coords),
                                   _ptr(np.ascontiguousarray(rotated_s1, dtype=np.float64)),
                                   _ptr(np.ascontiguousarray(xray, dtype=np.float64)),
                                   _ptr(np.ascontiguousarray(voxel_size, dtype=np.float64)),
                                   _ptr(np.ascontiguousarray(coefficients, dtype=np.float64)),
                                   _ptr(vol), _ptr(shape), int(full_iteration), int(is_exp),
                                   _ptr(rr) if rays else None, ct.byref(steps))
        return (m, rr.reshape(len(coords), 2)) if rays else m

    def cube_face(self, origin, direction, shape, L1):
        o = np.ascontiguousarray(origin, dtype=np.int64)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        s = np.ascontiguousarray(shape, dtype=np.int64)
        return int(self.lib.oracle_cube_face(_ptr(o), _ptr(d), _ptr(s), int(L1)))

    def thetaphi(self, v, L1):
        v = np.ascontiguousarray(v, dtype=np.float64)
        t, p = ct.c_double(), ct.c_double()
        self.lib.oracle_thetaphi(_ptr(v), int(L1), ct.byref(t), ct.byref(p))
        return t.value, p.value

    def face_rule(self, theta, phi, face):
        r = OracleRule()
        self.lib.oracle_face_rule(float(theta), float(phi), int(face), ct.byref(r))
        return dict(d=tuple(r.d), rounding=r.rounding, axis=r.axis, to_max=r.to_max,
                    inclusive=r.inclusive, valid=r.valid, vacuum_run=r.vacuum_run)

    def reflection_setup(self, s1, omega, raw_xray, omega_axis, F):
        s1 = np.ascontiguousarray(s1, dtype=np.float64)
        om = np.ascontiguousarray(omega, dtype=np.float64)
        out = np.zeros(len(om), dtype=REFL_DTYPE)
        self.lib.oracle_reflection_setup(_ptr(s1), _ptr(om), len(om),
                                         _ptr(np.ascontiguousarray(raw_xray, dtype=np.float64)),
                                         _ptr(np.ascontiguousarray(omega_axis, dtype=np.float64)),
                                         _ptr(np.ascontiguousarray(F, dtype=np.float64).reshape(-1)),
                                         _ptr(out))
        return out

    def kp_rotation(self, axis, theta):
        a = np.ascontiguousarray(axis, dtype=np.float64)
        out = np.zeros(9)
        self.lib.oracle_kp_rotation(_ptr(a), float(theta), _ptr(out))
        return out

    def single_ray(self, wl, coord, rotated_s1, theta, phi, is_exp: int = 0) -> float:
        """ray_tracing_single_gridding restated: one ray's sum of mu*L (or exp)."""
        vol = np.ascontiguousarray(wl.volume).view(np.uint8)
        shape = np.array(vol.shape, dtype=np.int64)
        return float(self.lib.oracle_single_ray(
            _ptr(np.ascontiguousarray(coord, dtype=np.int64)),
            _ptr(np.ascontiguousarray(rotated_s1, dtype=np.float64)), float(theta), float(phi),
            _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)), _ptr(vol),
            _ptr(shape), int(wl.full_iteration), int(is_exp), None))

    def gridding_map(self, wl, rotated_s1, theta_phi, is_exp: int = 0,
                     n_threads: int = 0) -> np.ndarray:
        """float32 [D][n_coords] absorption map of the workload's voxels."""
        vol = np.ascontiguousarray(wl.volume).view(np.uint8)
        shape = np.array(vol.shape, dtype=np.int64)
        coords = np.ascontiguousarray(wl.coords, dtype=np.int64)
        dirs = np.ascontiguousarray(rotated_s1, dtype=np.float64).reshape(-1, 3)
        tp = np.ascontiguousarray(theta_phi, dtype=np.float64).reshape(-1, 2)
        out = np.empty((len(dirs), len(coords)), dtype=np.float32)
        self.lib.oracle_gridding_map(
            _ptr(coords), len(coords), _ptr(dirs), _ptr(tp), len(dirs),
            _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
            _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)), _ptr(vol), _ptr(shape),
            int(wl.full_iteration), int(is_exp), int(n_threads or os.cpu_count() or 1), _ptr(out))
        return out


def label_table(volume: np.ndarray):
    """The int8*** row-pointer table that ``python_2_c_3d`` builds
    (utils/utils_os.py:45-68), built with numpy.  Returns (pointer, keepalive)."""
    vol = np.ascontiguousarray(volume)
    Z, Y, X = vol.shape
    base = vol.ctypes.data
    rows = (base + (np.arange(Z * Y, dtype=np.uint64) * np.uint64(X))).astype(np.uint64)
    planes = (rows.ctypes.data + np.arange(Z, dtype=np.uint64) * np.uint64(Y * 8)).astype(np.uint64)
    return ct.c_void_p(planes.ctypes.data), (vol, rows, planes)


@contextlib.contextmanager
def _stdout_silenced():
    """Redirect fd 1 (the reference prints one line per reflection)."""
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    libc = ct.CDLL(None)
    try:
        os.dup2(devnull, 1)
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


class Reference:
    """The reference's own C path built from its sources (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing; run `make -C oracle ref`")
        lib = ct.CDLL(path)
        lib.ray_tracing_overall.restype = ct.POINTER(ct.c_double)
        lib.ray_tracing_overall.argtypes = [ct.c_int64, ct.c_int64, ct.c_void_p, ct.c_int64,
                                            ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                            ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                                            ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int,
                                            ct.c_int]
        lib.free.restype = None
        lib.free.argtypes = [ct.c_void_p]
        lib.ray_tracing_single_gridding.restype = ct.c_double
        lib.ray_tracing_single_gridding.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_double,
                                                    ct.c_double, ct.c_void_p, ct.c_void_p,
                                                    ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int,
                                                    ct.c_int]
        lib.ib_test.restype = ct.c_double
        lib.ib_test.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int64,
                                ct.c_int]
        self.lib = lib

    def ib_test(self, wl, coords, rotated_s1, xray, num_cls: int = 3, table=None) -> float:
        """The reference's ib_test (ray_tracing.c:25-160).  It reads labels
        without bounds checks and asserts on inconsistent angles: call it in a
        process that may die."""
        keep = None
        if table is None:
            table, keep = label_table(np.ascontiguousarray(wl.volume).view(np.int8))
        shape = np.array(np.asarray(wl.volume).shape, dtype=np.int64)
        coords = np.ascontiguousarray(coords, dtype=np.int64)
        # ib_test prints nothing (its test_mode is 0), so unlike the other
        # entries it runs with stdout as it is: callable from several threads
        v = self.lib.ib_test(_ptr(coords), len(coords),
                             _ptr(np.ascontiguousarray(rotated_s1, dtype=np.float64)),
                             _ptr(np.ascontiguousarray(xray, dtype=np.float64)),
                             _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
                             _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)),
                             table, _ptr(shape), int(wl.full_iteration), 0, int(num_cls))
        del keep
        return float(v)

    def single_gridding(self, wl, coord, rotated_s1, theta, phi, is_exp: int = 0,
                        table=None) -> float:
        """The reference's ray_tracing_single_gridding (ray_tracing.c:484-534)."""
        keep = None
        if table is None:
            table, keep = label_table(np.ascontiguousarray(wl.volume).view(np.int8))
        shape = np.array(np.asarray(wl.volume).shape, dtype=np.int64)
        with _stdout_silenced():
            v = self.lib.ray_tracing_single_gridding(
                _ptr(np.ascontiguousarray(coord, dtype=np.int64)),
                _ptr(np.ascontiguousarray(rotated_s1, dtype=np.float64)), float(theta), float(phi),
                _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
                _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)), table, _ptr(shape),
                int(wl.full_iteration), 0, int(is_exp))
        del keep
        return float(v)

    def overall(self, wl, n_threads: int = 0, R: int | None = None) -> np.ndarray:
        s1 = np.ascontiguousarray(wl.s1, dtype=np.float64)
        om = np.ascontiguousarray(wl.omega, dtype=np.float64)
        if R is not None:
            s1, om = s1[:R], om[:R]
        coords = np.ascontiguousarray(wl.coords, dtype=np.int64)
        vol8 = np.ascontiguousarray(wl.volume).view(np.int8)
        table, keep = label_table(vol8)
        shape = np.array(vol8.shape, dtype=np.int64)
        nt = n_threads or os.cpu_count() or 1
        args = [0, len(om), _ptr(coords), len(coords), _ptr(s1), _ptr(om),
                _ptr(np.ascontiguousarray(wl.raw_xray, dtype=np.float64)),
                _ptr(np.ascontiguousarray(wl.omega_axis, dtype=np.float64)),
                _ptr(np.ascontiguousarray(wl.F, dtype=np.float64).reshape(-1)), len(om),
                _ptr(np.ascontiguousarray(wl.voxel_size, dtype=np.float64)),
                _ptr(np.ascontiguousarray(wl.coefficients, dtype=np.float64)),
                table, _ptr(shape), int(wl.full_iteration), 0, nt, 1]
        with _stdout_silenced():
            p = self.lib.ray_tracing_overall(*args)
        out = np.ctypeslib.as_array(p, shape=(len(om),)).copy()
        self.lib.free(p)
        del keep
        return out
