"""Loader and ctypes prototypes of libanacor_mi355x.so (include/anacor_mi355x.h).

The library is built in-tree by ``python -m anacor_public_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: if the library is missing
``load()`` raises, so nothing silently runs a CPU path in its place.
"""
from __future__ import annotations

import ctypes as ct
import os
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# ANACOR_LIB selects another build of the same library (e.g. the host-sanitizer
# build of `python -m anacor_public_amd.build --sanitize`)
LIB_PATH = os.environ.get("ANACOR_LIB", os.path.join(PKG, "lib", "libanacor_mi355x.so"))

ANACOR_OK = 0
ANACOR_ERR_ARG = 1
ANACOR_ERR_NO_DEVICE = 2
ANACOR_ERR_HIP = 3
ANACOR_ERR_GEOMETRY = 4
FLAG_F64_MEAN = 1
FLAG_NO_EXP = 2
LAYOUT_AUTO = -1
LAYOUT_RAW = 0
LAYOUT_BRICK = 1
LAYOUT_BRICK2 = 2
LAYOUT_BRICK4 = 4
RAY_ERR_FACE = 1
RAY_ERR_ASSERT = 2
NOALLOC_MAX_REFLECTIONS = 16384
NOALLOC_MAX_PARTIALS = 1 << 21
NOALLOC_SCRATCH_BYTES = 25428224
GATHER_LANES = 0
GATHER_ROW = 1

# Every symbol the header declares (checked by the CPU test suite).
EXPORTED = (
    "ray_tracing_overall", "ray_tracing_gpu_overall", "ray_tracing_single",
    "ray_tracing_single_mp", "anacor_last_error", "anacor_device_count", "anacor_ctx_create",
    "anacor_ctx_destroy", "anacor_ctx_set_tile", "anacor_ctx_set_order",
    "anacor_ctx_get_work_shape",
    "anacor_ctx_set_voxel_block",
    "anacor_ctx_set_layout", "anacor_ctx_get_layout", "anacor_ctx_run",
    "anacor_ctx_run_noalloc", "anacor_ctx_paths",
    "anacor_ctx_stats", "anacor_reflection_geometry", "anacor_batch_create",
    "anacor_batch_destroy", "anacor_batch_sample", "anacor_batch_run", "anacor_batch_stats",
    "anacor_batch_set_graph", "anacor_batch_last_graph",
    "anacor_mlp_create", "anacor_mlp_destroy", "anacor_mlp_num_params", "anacor_mlp_get_params",
    "anacor_mlp_set_params", "anacor_mlp_set_scale", "anacor_mlp_set_features",
    "anacor_mlp_set_envelope", "anacor_mlp_set_bands", "anacor_mlp_set_risk", "anacor_mlp_risk",
    "anacor_risk_fit", "anacor_mlp_features",
    "anacor_mlp_forward",
    "anacor_mlp_gradients", "anacor_mlp_train", "anacor_mlp_predict", "anacor_mlp_stats",
    "ray_tracing_single_gridding", "nearest_neighbor_interpolate_overall", "anacor_grid_create",
    "anacor_grid_destroy", "anacor_grid_map", "anacor_grid_run", "anacor_grid_stats",
    "ib_test", "anacor_ctx_bisection", "anacor_ctx_bisection_rotated", "anacor_probe_gather",
)

MLP_INPUTS = 60
MLP_WIDTH = 256
MLP_PARAMS = 213249
MLP_FEATURES_HARMONIC = 0
MLP_FEATURES_CHAMBER = 1
MLP_FEATURES_RESONANCE = 2

PATH_DTYPE = np.dtype([("face", np.int32, 2), ("len", np.int32, 2), ("nseg", np.int32, 2),
                       ("err", np.int32), ("pad", np.int32), ("last", np.int32, (2, 3)),
                       ("L", np.float64, (2, 4))])
assert PATH_DTYPE.itemsize == 120


class Stats(ct.Structure):
    _fields_ = [("kernel_ms", ct.c_double), ("voxel_steps", ct.c_uint64),
                ("workgroups", ct.c_int64), ("n_voxels", ct.c_int64), ("host_ms", ct.c_double),
                ("outside_reads", ct.c_uint64)]


class Sample(ct.Structure):
    _fields_ = [("volume", ct.c_void_p), ("shape", ct.c_int64 * 3), ("coords", ct.c_void_p),
                ("n_coords", ct.c_int64), ("voxel_size", ct.c_double * 3),
                ("coefficients", ct.c_double * 4)]


class Envelope(ct.Structure):
    _fields_ = [("axis", ct.c_double * 3), ("e1", ct.c_double * 3), ("e2", ct.c_double * 3),
                ("axial", ct.c_double * 2), ("two_theta", ct.c_double * 2),
                ("cells", ct.c_int32 * 3), ("pad", ct.c_int32)]


class Band(ct.Structure):
    _fields_ = [("ray", ct.c_int32), ("i", ct.c_int32), ("j", ct.c_int32), ("pad", ct.c_int32),
                ("ratio", ct.c_double), ("width", ct.c_double)]


MLP_MAX_BANDS = 32
RISK_FEATURES = 60
RISK_MAX_DEPTH = 8
RISK_MAX_TREES = 1024


class MlpStats(ct.Structure):
    _fields_ = [("mlp_ms", ct.c_double), ("total_ms", ct.c_double), ("rows", ct.c_int64),
                ("in_envelope", ct.c_int64)]


class AnacorError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"anacor error {code}: {message}")
        self.code = code


_lock = threading.Lock()
_lib = None


def _prototypes(lib):
    vp, i64, i32, dbl = ct.c_void_p, ct.c_int64, ct.c_int, ct.c_double
    lib.anacor_last_error.restype = ct.c_int
    lib.anacor_last_error.argtypes = [ct.c_char_p, ct.c_size_t]
    lib.anacor_device_count.restype = ct.c_int
    lib.anacor_device_count.argtypes = []
    lib.anacor_ctx_create.restype = vp
    lib.anacor_ctx_create.argtypes = [vp, vp, vp, i64, vp, vp, i32, i32]
    lib.anacor_ctx_destroy.restype = None
    lib.anacor_ctx_destroy.argtypes = [vp]
    lib.anacor_ctx_set_tile.restype = ct.c_int
    lib.anacor_ctx_set_tile.argtypes = [vp, i32]
    lib.anacor_ctx_get_work_shape.restype = ct.c_int
    lib.anacor_ctx_get_work_shape.argtypes = [vp, vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                              ct.POINTER(ct.c_int)]
    lib.anacor_ctx_set_order.restype = ct.c_int
    lib.anacor_ctx_set_order.argtypes = [vp, i32, i32]
    lib.anacor_ctx_set_layout.restype = ct.c_int
    lib.anacor_ctx_set_layout.argtypes = [vp, i32]
    lib.anacor_ctx_get_layout.restype = ct.c_int
    lib.anacor_ctx_get_layout.argtypes = [vp]
    lib.anacor_ctx_set_voxel_block.restype = ct.c_int
    lib.anacor_ctx_set_voxel_block.argtypes = [vp, i32, i32, i32]
    lib.anacor_ctx_run.restype = ct.c_int
    lib.anacor_ctx_run.argtypes = [vp, vp, vp, i64, vp, vp, vp, i32, vp]
    lib.anacor_ctx_run_noalloc.restype = ct.c_int
    lib.anacor_ctx_run_noalloc.argtypes = [vp, vp, vp, i64, vp, vp, vp, i32, vp]
    lib.anacor_ctx_paths.restype = ct.c_int
    lib.anacor_ctx_paths.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp]
    lib.anacor_ctx_stats.restype = ct.c_int
    lib.anacor_ctx_stats.argtypes = [vp, ct.POINTER(Stats)]
    lib.anacor_batch_create.restype = vp
    lib.anacor_batch_create.argtypes = [ct.POINTER(Sample), i32, i32, i32]
    lib.anacor_batch_destroy.restype = None
    lib.anacor_batch_destroy.argtypes = [vp]
    lib.anacor_batch_sample.restype = vp
    lib.anacor_batch_sample.argtypes = [vp, i32]
    lib.anacor_batch_run.restype = ct.c_int
    lib.anacor_batch_run.argtypes = [vp, vp, vp, vp, vp, vp, vp, i32, vp]
    lib.anacor_batch_set_graph.restype = ct.c_int
    lib.anacor_batch_set_graph.argtypes = [vp, i32]
    lib.anacor_batch_last_graph.restype = ct.c_int
    lib.anacor_batch_last_graph.argtypes = [vp]
    lib.anacor_batch_stats.restype = ct.c_int
    lib.anacor_batch_stats.argtypes = [vp, ct.POINTER(Stats)]
    lib.anacor_reflection_geometry.restype = ct.c_int
    lib.anacor_reflection_geometry.argtypes = [vp, vp, i64, vp, vp, vp, i64, vp, vp]
    u64, f32, sz = ct.c_uint64, ct.c_float, ct.c_size_t
    lib.anacor_mlp_create.restype = vp
    lib.anacor_mlp_create.argtypes = [u64, i32]
    lib.anacor_mlp_destroy.restype = None
    lib.anacor_mlp_destroy.argtypes = [vp]
    lib.anacor_mlp_num_params.restype = sz
    lib.anacor_mlp_num_params.argtypes = []
    lib.anacor_mlp_get_params.restype = ct.c_int
    lib.anacor_mlp_get_params.argtypes = [vp, vp, sz]
    lib.anacor_mlp_set_params.restype = ct.c_int
    lib.anacor_mlp_set_params.argtypes = [vp, vp, sz]
    lib.anacor_mlp_set_scale.restype = ct.c_int
    lib.anacor_mlp_set_scale.argtypes = [vp, dbl, dbl]
    lib.anacor_mlp_set_features.restype = ct.c_int
    lib.anacor_mlp_set_features.argtypes = [vp, i32]
    lib.anacor_mlp_set_envelope.restype = ct.c_int
    lib.anacor_mlp_set_envelope.argtypes = [vp, ct.POINTER(Envelope), vp]
    lib.anacor_mlp_set_bands.restype = ct.c_int
    lib.anacor_mlp_set_bands.argtypes = [vp, ct.POINTER(Band), i32]
    lib.anacor_mlp_set_risk.restype = ct.c_int
    lib.anacor_mlp_set_risk.argtypes = [vp, i32, i32, vp, vp, vp, ct.c_float, ct.c_float]
    lib.anacor_mlp_risk.restype = ct.c_int
    lib.anacor_mlp_risk.argtypes = [vp, vp, i64, vp, vp]
    lib.anacor_risk_fit.restype = ct.c_int
    lib.anacor_risk_fit.argtypes = [vp, vp, vp, i64, i32, i32, i32, ct.c_float, ct.c_float,
                                    ct.c_float, i32, i32, vp, vp, vp, ct.POINTER(ct.c_float)]
    lib.anacor_mlp_features.restype = ct.c_int
    lib.anacor_mlp_features.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, vp]
    lib.anacor_mlp_forward.restype = ct.c_int
    lib.anacor_mlp_forward.argtypes = [vp, vp, i64, vp]
    lib.anacor_mlp_gradients.restype = ct.c_int
    lib.anacor_mlp_gradients.argtypes = [vp, vp, vp, i64, vp, vp]
    lib.anacor_mlp_train.restype = ct.c_int
    lib.anacor_mlp_train.argtypes = [vp, vp, vp, i64, i32, i32, f32, f32, u64, vp]
    lib.anacor_mlp_predict.restype = ct.c_int
    lib.anacor_mlp_predict.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp, vp]
    lib.anacor_mlp_stats.restype = ct.c_int
    lib.anacor_mlp_stats.argtypes = [vp, ct.POINTER(MlpStats)]
    lib.anacor_grid_create.restype = vp
    lib.anacor_grid_create.argtypes = [vp, i64, i64, vp, vp, i32]
    lib.anacor_grid_destroy.restype = None
    lib.anacor_grid_destroy.argtypes = [vp]
    lib.anacor_grid_map.restype = ct.c_int
    lib.anacor_grid_map.argtypes = [vp, vp]
    lib.anacor_grid_run.restype = ct.c_int
    lib.anacor_grid_run.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp, i64, vp, i64, i32, vp]
    lib.anacor_grid_stats.restype = ct.c_int
    lib.anacor_grid_stats.argtypes = [vp, ct.POINTER(dbl), ct.POINTER(dbl)]
    # gridding drop-ins (utils/utils_gridding.py:104-118, :567-591)
    lib.ray_tracing_single_gridding.restype = dbl
    lib.ray_tracing_single_gridding.argtypes = [vp, vp, dbl, dbl, vp, vp, vp, vp, i32, i32, i32]
    lib.nearest_neighbor_interpolate_overall.restype = ct.POINTER(dbl)
    lib.nearest_neighbor_interpolate_overall.argtypes = [
        i64, i64, vp, i64, vp, vp, vp, vp, vp, i64, vp, vp, i32, i32, vp, vp, vp, ct.c_size_t,
        ct.c_size_t, dbl, dbl, dbl, dbl, i32]
    # drop-in entry points, with the argtypes the reference's binding uses
    # (utils/utils_mp.py:59-79, :82-94, :128-148)
    lib.ray_tracing_overall.restype = ct.POINTER(dbl)
    lib.ray_tracing_overall.argtypes = [i64, i64, vp, i64, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp,
                                        i32, i32, i32, i32]
    lib.ray_tracing_gpu_overall.restype = ct.POINTER(ct.c_float)
    lib.ray_tracing_gpu_overall.argtypes = [i64, i64, vp, i64, vp, vp, vp, vp, vp, i64, vp, vp, vp,
                                            vp, vp, i32, i32, i32]
    lib.ray_tracing_single.restype = dbl
    lib.ray_tracing_single.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, i32, i64, i32]
    lib.ray_tracing_single_mp.restype = dbl
    lib.ray_tracing_single_mp.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, i32, i64, i32, i32]
    # iterative bisection (binding mp_worker.py:390-404; its trailing IsExp
    # argument is dropped here, the callee ignores it)
    lib.ib_test.restype = dbl
    lib.ib_test.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, i32, i64, i32]
    lib.anacor_ctx_bisection.restype = ct.c_int
    lib.anacor_ctx_bisection.argtypes = [vp, vp, vp, i64, vp, vp, vp, i32, i32, vp]
    lib.anacor_ctx_bisection_rotated.restype = ct.c_int
    lib.anacor_ctx_bisection_rotated.argtypes = [vp, vp, vp, i64, i32, i32, vp, vp]
    lib.anacor_probe_gather.restype = ct.c_int
    lib.anacor_probe_gather.argtypes = [i32, i32, i32, i64, ct.POINTER(ct.c_double)]
    lib.free.restype = None
    lib.free.argtypes = [vp]
    return lib


def load(path: str = LIB_PATH):
    """Load (once) and return the library; raise if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise FileNotFoundError(
                    f"{path} is missing: build it with `python -m anacor_public_amd.build`")
            _lib = _prototypes(ct.CDLL(path))
        return _lib


def last_error(lib=None):
    lib = lib or load()
    buf = ct.create_string_buffer(512)
    code = lib.anacor_last_error(buf, 512)
    return code, buf.value.decode(errors="replace")


def check(rc: int, lib=None, allow=(ANACOR_OK,)):
    if rc not in allow:
        code, msg = last_error(lib)
        raise AnacorError(rc, msg)
    return rc
