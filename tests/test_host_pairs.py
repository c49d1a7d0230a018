"""The reference's one-pair C functions (libsupermesh-c.h:24-31;
C_bindings.F90:42-64,134-174) run on the calling thread, with no device:
bitwise equal to the oracle on the reference's fixture meshes and jittered
grids, re-entrant across threads.  No GPU is needed for any of this."""
import ctypes
import threading

import numpy as np
import pytest

import libsupermesh_amd
import oracle_lib as O
from libsupermesh_amd import meshgen

_dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def L():
    lib = ctypes.CDLL(libsupermesh_amd.get_library())
    for name in ("libsupermesh_intersect_tris_real", "libsupermesh_intersect_tets_real",
                 "libsupermesh_intersect_tri_quad", "libsupermesh_intersect_tet_hex",
                 "libsupermesh_intersect_quads", "libsupermesh_intersect_hexes"):
        getattr(lib, name).argtypes = [_dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
    lib.libsupermesh_triangle_area.argtypes = [_dp, _dp]
    lib.libsupermesh_tetrahedron_volume.argtypes = [_dp, _dp]
    return lib


def _call(fn, a, b, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(shape)
    n = ctypes.c_int(-1)
    fn(a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), out.ctypes.data_as(_dp), ctypes.byref(n))
    return out[:n.value].copy()


def _pairs(pa, ea, pb, eb, closed=False, limit=None):
    ip, ix = O.brute_force(pa, ea, pb, eb, closed=closed)
    rows, cols = O.csr_to_pairs(ip, ix)
    if limit is not None and len(rows) > limit:
        sel = np.random.default_rng(5).choice(len(rows), limit, replace=False)
        rows, cols = rows[sel], cols[sel]
    return rows, cols


@pytest.mark.parametrize("names", [("triangle.1", "triangle.2"), ("cube.1", "cube.2"), ("tet", "cube.2")])
def test_simplex_pairs_bitwise(L, fixture_meshes, names):
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    dim = pa.shape[1]
    rows, cols = _pairs(pa, ea, pb, eb, limit=4000)
    fn = L.libsupermesh_intersect_tris_real if dim == 2 else L.libsupermesh_intersect_tets_real
    shape = (22, 3, 2) if dim == 2 else (81, 4, 3)
    ref = O.intersect_tris if dim == 2 else O.intersect_tets
    nonempty = 0
    for r, c in zip(rows, cols):
        got = _call(fn, pa[ea[r]], pb[eb[c]], shape)
        exp = ref(pa[ea[r]], pb[eb[c]])
        assert got.shape == exp.shape and np.array_equal(got, exp), (r, c)
        nonempty += len(got) > 0
    assert nonempty > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_jittered_grid_pairs_bitwise(L, dim):
    pa, ea = meshgen.split_grid(dim, 10 if dim == 2 else 4, 0)
    pb, eb = meshgen.split_grid(dim, 9 if dim == 2 else 3, 1, mesh_id=1, diagonal=1)
    rows, cols = _pairs(pa, ea, pb, eb)
    fn = L.libsupermesh_intersect_tris_real if dim == 2 else L.libsupermesh_intersect_tets_real
    shape = (22, 3, 2) if dim == 2 else (81, 4, 3)
    cnt, meas = O.clip_pairs(pa, ea, pb, eb, rows, cols)
    measure = L.libsupermesh_triangle_area if dim == 2 else L.libsupermesh_tetrahedron_volume
    total = 0.0
    for i, (r, c) in enumerate(zip(rows, cols)):
        got = _call(fn, pa[ea[r]], pb[eb[c]], shape)
        assert len(got) == cnt[i]
        m = 0.0
        for t in got:
            v = ctypes.c_double()
            measure(np.ascontiguousarray(t).ctypes.data_as(_dp), ctypes.byref(v))
            m += v.value
        assert m == meas[i]
        total += m
    assert abs(total - 1.0) < 1e-12


@pytest.mark.parametrize("names,kind", [(("square.1", "square.4"), "tri_quad"),
                                        (("square.3", "square.4"), "quads"),
                                        (("cube.1", "cube.3"), "tet_hex"),
                                        (("cube.3", "cube.4"), "hexes")])
def test_mixed_pairs_bitwise(L, fixture_meshes, names, kind):
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    rows, cols = _pairs(pa, ea, pb, eb, closed=True, limit=1500)
    fn, shape, ref = {
        "tri_quad": (L.libsupermesh_intersect_tri_quad, (46, 3, 2), O.intersect_tri_quad),
        "quads": (L.libsupermesh_intersect_quads, (62, 3, 2), O.intersect_quads),
        "tet_hex": (L.libsupermesh_intersect_tet_hex, (729, 4, 3), O.intersect_tet_hex),
        "hexes": (L.libsupermesh_intersect_hexes, (5 * 729, 4, 3), O.intersect_hexes),
    }[kind]
    for r, c in zip(rows, cols):
        got = _call(fn, pa[ea[r]], pb[eb[c]], shape)
        assert np.array_equal(got, ref(pa[ea[r]], pb[eb[c]])), (r, c)


def test_measures_bitwise(L):
    rng = np.random.default_rng(3)
    for _ in range(200):
        t = rng.standard_normal((3, 2))
        v = ctypes.c_double()
        L.libsupermesh_triangle_area(np.ascontiguousarray(t).ctypes.data_as(_dp), ctypes.byref(v))
        assert v.value == O.triangle_area(t)
        T = rng.standard_normal((4, 3))
        L.libsupermesh_tetrahedron_volume(np.ascontiguousarray(T).ctypes.data_as(_dp), ctypes.byref(v))
        assert v.value == O.tet_volume_real(T)


def test_reentrant_across_threads(L):
    """Eight threads clipping different pairs at once get the serial results."""
    pa, ea = meshgen.split_grid(3, 4, 0)
    pb, eb = meshgen.split_grid(3, 3, 1, mesh_id=1, diagonal=1)
    rows, cols = _pairs(pa, ea, pb, eb)
    expect = [O.intersect_tets(pa[ea[r]], pb[eb[c]]) for r, c in zip(rows, cols)]
    bad = []

    def work(t):
        for i in range(t, len(rows), 8):
            got = _call(L.libsupermesh_intersect_tets_real, pa[ea[rows[i]]], pb[eb[cols[i]]], (81, 4, 3))
            if not np.array_equal(got, expect[i]):
                bad.append(i)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not bad


def test_success_clears_the_last_error(L):
    """Like every other call, a per-pair call that succeeds clears the thread's
    last error (a failed call before it leaves one)."""
    L.libsupermesh_gpu_last_error.restype = ctypes.c_char_p
    n = ctypes.c_long(0)
    L.libsupermesh_sort_intersection_finder_query_output(ctypes.byref(n))   # before set_input
    assert L.libsupermesh_gpu_last_error_code() == 3
    _call(L.libsupermesh_intersect_tris_real, [[0, 0], [1, 0], [0, 1]], [[0, 0], [1, 0], [0, 1]], (22, 3, 2))
    assert L.libsupermesh_gpu_last_error_code() == 0 and L.libsupermesh_gpu_last_error() == b""
