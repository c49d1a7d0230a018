"""libsupermesh_gpu_clip_pairs_noalloc against libsupermesh_gpu_clip_pairs,
through the C ABI: the same per-pair counts and measures (bitwise) and the
same measure sum, on two calls in a row, the second with a smaller list; the
size limit and out-of-range ids."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from libsupermesh_amd import gpu, meshgen

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 22


def _lib():
    L = gpu.lib()
    f = L.libsupermesh_gpu_clip_pairs_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(gpu.Mesh), ctypes.POINTER(gpu.Mesh), ctypes.c_long,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return L


def _noalloc(A, B, pa, pb, n, counts, measure, total):
    L = _lib()
    rc = L.libsupermesh_gpu_clip_pairs_noalloc(0, None, ctypes.byref(A.struct), ctypes.byref(B.struct), n,
                                              pa.ptr, pb.ptr, counts.ptr, measure.ptr, total.ptr)
    gpu._check(rc)
    gpu.synchronize()
    return (gpu.device_to_numpy(counts.ptr, n, np.int32), gpu.device_to_numpy(measure.ptr, n, np.float64),
            float(gpu.device_to_numpy(total.ptr, 1, np.float64)[0]))


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_equals_clip_pairs_twice(dim):
    if dim == 2:
        pa_, ea = meshgen.split_grid(2, 40, 1)
        pb_, eb = meshgen.split_grid(2, 37, 2, mesh_id=1, diagonal=1)
    else:
        pa_, ea = meshgen.split_grid(3, 8, 1)
        pb_, eb = meshgen.split_grid(3, 7, 2, mesh_id=1, diagonal=1)
    rows, cols = O.csr_to_pairs(*O.brute_force(pa_, ea, pb_, eb))
    A, B = gpu.DeviceMesh(pa_, ea), gpu.DeviceMesh(pb_, eb)
    ctx = gpu.Context()
    n = rows.shape[0]
    counts = gpu.DeviceArray(np.zeros(n, np.int32))
    measure = gpu.DeviceArray(np.zeros(n, np.float64))
    total = gpu.DeviceArray(np.zeros(1, np.float64))
    for m in (n, n // 3):   # the second call smaller: scratch keeps the first call's contents
        pa = gpu.DeviceArray(np.ascontiguousarray(rows[:m]))
        pb = gpu.DeviceArray(np.ascontiguousarray(cols[:m]))
        r = ctx.clip_pairs(A, B, pa, pb)
        ref = gpu.fetch(r, dim, ("counts", "measure"))
        c, meas, s = _noalloc(A, B, pa, pb, m, counts, measure, total)
        assert np.array_equal(c, ref["counts"])
        assert np.array_equal(meas, ref["measure"])
        assert s == r.sum
    # and against the oracle
    cnt, ms = O.clip_pairs(pa_, ea, pb_, eb, rows[: n // 3], cols[: n // 3])
    assert np.array_equal(c, cnt) and np.array_equal(meas, ms)
    ctx.close()


def test_noalloc_limits():
    p, e = meshgen.split_grid(2, 4, 0)
    A = gpu.DeviceMesh(p, e)
    L = _lib()
    buf = gpu.DeviceArray(np.zeros(8, np.int32))
    rc = L.libsupermesh_gpu_clip_pairs_noalloc(0, None, ctypes.byref(A.struct), ctypes.byref(A.struct),
                                              MAX_PAIRS + 1, buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr)
    assert rc == 1 and b"4194304" in L.libsupermesh_gpu_last_error()
    # ids outside the meshes: count -1, measure 0, no read through them
    pa = gpu.DeviceArray(np.array([0, e.shape[0], 1, -1], np.int32))
    pb = gpu.DeviceArray(np.array([0, 0, e.shape[0] + 7, 2], np.int32))
    counts = gpu.DeviceArray(np.zeros(4, np.int32))
    measure = gpu.DeviceArray(np.zeros(4, np.float64))
    total = gpu.DeviceArray(np.zeros(1, np.float64))
    c, m, s = _noalloc(A, A, pa, pb, 4, counts, measure, total)
    assert c[0] == 1 and list(c[1:]) == [-1, -1, -1]
    assert m[0] == O.triangle_area(p[e[0]]) and not np.any(m[1:])
    assert s == m[0]
