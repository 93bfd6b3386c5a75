"""The drop-in libraries loaded and called exactly as the reference's driver
does it: the file names utils/utils_mp.py:51-55 and :119-125 load
(ray_tracing_cpu.so, ray_tracing_gpu_{card}.so with --gpu-card mi355x), the
literal argtypes of utils_mp.py:59-79 (ray_tracing_overall), :81-94
(ray_tracing_single, with c_int for its two int64_t parameters), :95-109
(ib_test, with its extra trailing IsExp) and :128-148
(ray_tracing_gpu_overall), the int8*** table built with ctypes arrays as
python_2_c_3d does (utils/utils_os.py:45-68, called at utils_mp.py:110), the
calls of :178-187 and :197-204 and the release of the result with the
library's libc free (:212).

Expected values: the reference build's own A_h for the 64^3 golden case
(tests/golden/golden_small.npz, FP32 exp and compensated mean on the GPU, so
1e-6 relative); the FP32 entry against the oracle on the float32-rounded
inputs widened to double (2e-6)."""
import ctypes as ct
import dataclasses
import os

import numpy as np
import pytest

from conftest import workload_from

pytestmark = pytest.mark.gpu

REL = 1e-6
LIBDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "anacor_public_amd", "lib")
i8p = ct.POINTER(ct.c_int8)
i8pp = ct.POINTER(i8p)
i8ppp = ct.POINTER(i8pp)
nd = np.ctypeslib.ndpointer


def row_pointer_table(label_list):
    """The int8*** table of utils_os.py:45-68: per z-plane a ctypes array of
    row pointers into the numpy array, and an array of those planes."""
    planes = []
    for plane in label_list:
        rows = (i8p * label_list.shape[1])(*[row.ctypes.data_as(i8p) for row in plane])
        planes.append(ct.cast(rows, i8pp))
    cube = (i8pp * label_list.shape[0])(*planes)
    return ct.cast(cube, i8ppp)


def bind_cpu(lib):
    lib.ray_tracing_overall.restype = ct.POINTER(ct.c_double)
    lib.ray_tracing_overall.argtypes = [
        ct.c_int64, ct.c_int64, nd(dtype=np.int64), ct.c_int64, nd(dtype=np.float64),
        nd(dtype=np.float64), nd(dtype=np.float64), nd(dtype=np.float64), nd(dtype=np.float64),
        ct.c_int64, nd(dtype=np.float64), nd(dtype=np.float64), i8ppp, nd(dtype=np.int64),
        ct.c_int, ct.c_int, ct.c_int, ct.c_int]
    lib.ray_tracing_single.restype = ct.c_double
    lib.ray_tracing_single.argtypes = [
        nd(dtype=np.int64), ct.c_int, nd(dtype=np.float64), nd(dtype=np.float64),
        nd(dtype=np.float64), nd(dtype=np.float64), i8ppp, nd(dtype=np.int64), ct.c_int,
        ct.c_int, ct.c_int]
    lib.ib_test.restype = ct.c_double
    lib.ib_test.argtypes = [
        nd(dtype=np.int64), ct.c_int, nd(dtype=np.float64), nd(dtype=np.float64),
        nd(dtype=np.float64), nd(dtype=np.float64), i8ppp, nd(dtype=np.int64), ct.c_int,
        ct.c_int, ct.c_int, ct.c_int]
    return lib


def bind_gpu(lib):
    lib.ray_tracing_gpu_overall.restype = ct.POINTER(ct.c_float)
    lib.ray_tracing_gpu_overall.argtypes = [
        ct.c_int64, ct.c_int64, nd(dtype=np.int32), ct.c_int64, nd(dtype=np.float32),
        nd(dtype=np.float32), nd(dtype=np.float32), nd(dtype=np.float32), nd(dtype=np.float32),
        ct.c_int64, nd(dtype=np.float32), nd(dtype=np.float32), i8ppp, i8p,
        nd(dtype=np.int32), ct.c_int, ct.c_int, ct.c_int]
    return lib


@pytest.fixture(scope="module")
def libs(gpu):
    cpu = bind_cpu(ct.CDLL(os.path.join(LIBDIR, "ray_tracing_cpu.so")))
    gpu_lib = bind_gpu(ct.CDLL(os.path.join(LIBDIR, "ray_tracing_gpu_mi355x.so")))
    return cpu, gpu_lib


def test_overall_as_worker_function_calls_it(libs, golden_small):
    cpu, _ = libs
    wl = workload_from(golden_small)
    label_list = np.ascontiguousarray(wl.volume).astype(np.int8)   # main_lite.py:111
    label_list_c = row_pointer_table(label_list)
    shape = np.array(label_list.shape)
    low, R = 0, len(wl.omega)
    for full, key in ((0, "A_full0"), (1, "A_full1")):
        result_list = cpu.ray_tracing_overall(
            low, low + R, wl.coords, len(wl.coords), wl.s1, wl.omega, wl.raw_xray,
            wl.omega_axis, wl.F, R, wl.voxel_size, wl.coefficients, label_list_c, shape, full, 0,
            8, 1)
        corr = [result_list[i] for i in range(R)]
        cpu.free(result_list)
        np.testing.assert_allclose(corr, golden_small[key], rtol=REL, atol=0)


def test_gpu_overall_as_worker_function_calls_it(libs, oracle, golden_small):
    cpu, gpu_lib = libs
    wl = workload_from(golden_small).subset(16)
    label_list = np.ascontiguousarray(wl.volume).astype(np.int8)
    label_list_c = row_pointer_table(label_list)
    shape = np.array(label_list.shape)
    low, R = 0, len(wl.omega)
    result_list = gpu_lib.ray_tracing_gpu_overall(
        low, low + R, wl.coords.astype(np.int32), np.int64(len(wl.coords)),
        wl.s1.astype(np.float32), wl.omega.astype(np.float32), wl.raw_xray.astype(np.float32),
        wl.omega_axis.astype(np.float32), wl.F.astype(np.float32), np.int64(R),
        wl.voxel_size.astype(np.float32), wl.coefficients.astype(np.float32), label_list_c,
        label_list.ctypes.data_as(i8p), shape.astype(np.int32), 0, 0, 1)
    corr = np.array([result_list[i] for i in range(R)], dtype=np.float32)
    cpu.free(result_list)   # utils_mp.py:212 frees through the CPU library
    widened = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    w32 = dataclasses.replace(wl, s1=widened(wl.s1), omega=widened(wl.omega),
                              raw_xray=widened(wl.raw_xray), omega_axis=widened(wl.omega_axis),
                              F=widened(wl.F), voxel_size=widened(wl.voxel_size),
                              coefficients=widened(wl.coefficients))
    np.testing.assert_allclose(corr, oracle.trace(w32)["result"], rtol=2e-6, atol=0)


def test_single_and_ib_test_with_c_int_lengths(libs, oracle, golden_small):
    """ray_tracing_single and ib_test declare int64_t lengths; the binding
    passes c_int (and ib_test one argument more), as the reference's does."""
    cpu, _ = libs
    wl = workload_from(golden_small)
    label_list = np.ascontiguousarray(wl.volume).astype(np.int8)
    label_list_c = row_pointer_table(label_list)
    shape = np.array(label_list.shape)
    g = oracle.reflection_setup(wl.s1[:3], wl.omega[:3], wl.raw_xray, wl.omega_axis, wl.F)
    for i in range(3):
        s1r = np.ascontiguousarray(g["s1"][i])
        xr = np.ascontiguousarray(g["xray"][i])
        v = cpu.ray_tracing_single(wl.coords, len(wl.coords), s1r, xr, wl.voxel_size,
                                   wl.coefficients, label_list_c, shape, 0, 0, 1)
        assert abs(v - golden_small["A_full0"][i]) <= REL * golden_small["A_full0"][i]
        ib = cpu.ib_test(wl.coords, len(wl.coords), s1r, xr, wl.voxel_size, wl.coefficients,
                         label_list_c, shape, 0, 0, 3, 1)
        exp, _ = oracle.ib_test(wl, wl.coords, s1r, xr, 3)
        assert abs(ib - exp) <= REL * exp
