"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is the checker the GPU results are
compared against, never the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_llp = ctypes.POINTER(ctypes.c_longlong)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_spacing.restype = ctypes.c_double
        L.orc_spacing.argtypes = [ctypes.c_double]
        L.orc_norm2_3.restype = ctypes.c_double
        L.orc_norm2_3.argtypes = [_dp]
        L.orc_dot3.restype = ctypes.c_double
        L.orc_dot3.argtypes = [_dp, _dp]
        L.orc_triangle_area.restype = ctypes.c_double
        L.orc_triangle_area.argtypes = [_dp]
        L.orc_tetrahedron_volume_real.restype = ctypes.c_double
        L.orc_tetrahedron_volume_real.argtypes = [_dp]
        L.orc_tetrahedron_volume_tet.restype = ctypes.c_double
        L.orc_tetrahedron_volume_tet.argtypes = [_dp]
        L.orc_intersect_tris.restype = ctypes.c_int
        L.orc_intersect_tris.argtypes = [_dp, _dp, _dp]
        L.orc_intersect_tets.restype = ctypes.c_int
        L.orc_intersect_tets.argtypes = [_dp, _dp, _dp]
        L.orc_get_planes_tet.argtypes = [_dp, _dp, _dp]
        L.orc_brute_force_finder.restype = ctypes.c_long
        L.orc_brute_force_finder.argtypes = [ctypes.c_int, ctypes.c_int,
                                             ctypes.c_long, ctypes.c_int, _dp, _ip,
                                             ctypes.c_long, ctypes.c_int, _dp, _ip, _lp, _ip]
        L.orc_tree_finder.restype = ctypes.c_long
        L.orc_tree_finder.argtypes = [ctypes.c_int, ctypes.c_long,
                                      ctypes.c_long, ctypes.c_int, _dp, _ip,
                                      ctypes.c_long, ctypes.c_int, _dp, _ip, _lp, _ip]
        L.orc_clip_pairs.argtypes = [ctypes.c_int, ctypes.c_long, _ip, _ip, _dp, _ip, _dp, _ip,
                                     _ip, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_mass_pairs.argtypes = [ctypes.c_int, ctypes.c_long, _ip, _ip, _dp, _ip, _dp, _ip,
                                     _dp, ctypes.c_int]
        L.orc_inner_product.argtypes = [ctypes.c_int, ctypes.c_long, _ip, _ip, _dp, _ip, _dp,
                                        _dp, _ip, _dp, _dp, _dp]
        L.orc_supermesh.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_long, ctypes.c_long, _dp, _ip,
                                    ctypes.c_long, ctypes.c_long, _dp, _ip, _llp, _dp, _dp]
        for name in ("orc_intersect_tri_quad", "orc_intersect_quads", "orc_intersect_tet_hex",
                     "orc_intersect_hexes", "orc_intersect_intervals"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [_dp, _dp, _dp]
        L.orc_hex_planes.argtypes = [_dp, _dp, _dp]
        L.orc_interval_size.restype = ctypes.c_double
        L.orc_interval_size.argtypes = [_dp]
        L.orc_sort_finder.restype = ctypes.c_long
        L.orc_sort_finder.argtypes = [ctypes.c_long, _dp, _ip, ctypes.c_long, _dp, _ip, _lp, _ip,
                                      ctypes.c_long]
        _l = ctypes.c_long
        L.orc_mesh_nelist.argtypes = [_l, _l, ctypes.c_int, _lp, _lp, _lp]
        L.orc_mesh_nnlist.restype = _l
        L.orc_mesh_nnlist.argtypes = [_l, _l, ctypes.c_int, _lp, _lp, _lp]
        L.orc_mesh_eelist.restype = _l
        L.orc_mesh_eelist.argtypes = [_l, _l, ctypes.c_int, _l, ctypes.c_int, _lp, _lp, _lp]
        L.orc_transpose_sparsity.argtypes = [_l, _l, _lp, _lp, _lp, _lp]
        L.orc_advancing_front_finder.restype = _l
        L.orc_advancing_front_finder.argtypes = [ctypes.c_int, _l, _l, ctypes.c_int, _dp, _lp,
                                                 _l, _l, ctypes.c_int, _dp, _lp, _lp,
                                                 ctypes.POINTER(_lp)]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_grid_finder.restype = ctypes.c_long
        L.orc_grid_finder.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_long, ctypes.c_int, _dp, _ip,
                                      ctypes.c_long, ctypes.c_int, _dp, _ip,
                                      _lp, ctypes.POINTER(_ip)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def spacing(x):
    return lib().orc_spacing(float(x))


def norm2(v):
    v = _f64(v)
    return lib().orc_norm2_3(_d(v))


def dot3(a, b):
    a, b = _f64(a), _f64(b)
    return lib().orc_dot3(_d(a), _d(b))


def triangle_area(t):
    t = _f64(t)
    return lib().orc_triangle_area(_d(t))


def tet_volume_real(t):
    t = _f64(t)
    return lib().orc_tetrahedron_volume_real(_d(t))


def tet_volume_tet(t):
    t = _f64(t)
    return lib().orc_tetrahedron_volume_tet(_d(t))


def intersect_tris(a, b):
    a, b = _f64(a), _f64(b)
    out = np.zeros((22, 3, 2))
    n = lib().orc_intersect_tris(_d(a), _d(b), _d(out))
    return out[:n].copy()


def intersect_tets(a, b):
    a, b = _f64(a), _f64(b)
    out = np.zeros((81, 4, 3))
    n = lib().orc_intersect_tets(_d(a), _d(b), _d(out))
    return out[:n].copy()


def _mixed(fn, a, b, shape):
    a, b = _f64(a), _f64(b)
    out = np.zeros(shape)
    n = fn(_d(a), _d(b), _d(out))
    return out[:n].copy()


def intersect_tri_quad(tri, quad):
    return _mixed(lib().orc_intersect_tri_quad, tri, quad, (46, 3, 2))


def intersect_quads(qa, qb):
    return _mixed(lib().orc_intersect_quads, qa, qb, (62, 3, 2))


def intersect_tet_hex(tet, hex_):
    return _mixed(lib().orc_intersect_tet_hex, tet, hex_, (729, 4, 3))


def intersect_hexes(ha, hb):
    return _mixed(lib().orc_intersect_hexes, ha, hb, (5 * 729, 4, 3))


def hex_planes(h):
    h = _f64(h)
    n = np.zeros((6, 3))
    c = np.zeros(6)
    lib().orc_hex_planes(_d(h), _d(n), _d(c))
    return n, c


def interval_size(v):
    v = _f64(v)
    return lib().orc_interval_size(_d(v))


def intersect_intervals(a, b):
    """(n, interval) as intersect_intervals: n is 0 or 1."""
    a, b = _f64(a), _f64(b)
    c = np.zeros(2)
    n = lib().orc_intersect_intervals(_d(a), _d(b), _d(c))
    return n, c


def sort_finder(pos_a, enl_a, pos_b, enl_b):
    pos_a, pos_b = _f64(pos_a).reshape(-1), _f64(pos_b).reshape(-1)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    ne_a, ne_b = enl_a.shape[0], enl_b.shape[0]
    indptr = np.zeros(ne_a + 1, dtype=np.int64)
    L = lib()
    n = L.orc_sort_finder(ne_a, _d(pos_a), _i(enl_a), ne_b, _d(pos_b), _i(enl_b),
                          indptr.ctypes.data_as(_lp), None, 0)
    indices = np.zeros(max(n, 1), dtype=np.int32)
    L.orc_sort_finder(ne_a, _d(pos_a), _i(enl_a), ne_b, _d(pos_b), _i(enl_b),
                      indptr.ctypes.data_as(_lp), _i(indices), n)
    return indptr, indices[:n]


def get_planes_tet(t):
    t = _f64(t)
    n = np.zeros((4, 3))
    c = np.zeros(4)
    lib().orc_get_planes_tet(_d(t), _d(n), _d(c))
    return n, c


def brute_force(pos_a, enl_a, pos_b, enl_b, closed=False):
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    dim = pos_a.shape[1]
    ne_a, ne_b = enl_a.shape[0], enl_b.shape[0]
    indptr = np.zeros(ne_a + 1, dtype=np.int64)
    L = lib()
    args = (dim, int(closed), ne_a, enl_a.shape[1], _d(pos_a), _i(enl_a),
            ne_b, enl_b.shape[1], _d(pos_b), _i(enl_b))
    L.orc_brute_force_finder(*args, indptr.ctypes.data_as(_lp), None)
    indices = np.zeros(max(int(indptr[-1]), 1), dtype=np.int32)
    L.orc_brute_force_finder(*args, indptr.ctypes.data_as(_lp), _i(indices))
    return indptr, indices[:indptr[-1]]


def grid_finder(pos_a, enl_a, pos_b, enl_b, closed=False, nthreads=8):
    """The brute-force set (rows ascending) through a uniform grid of cells:
    the same predicate as brute_force, fast enough for the full workloads."""
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    ne_a = enl_a.shape[0]
    indptr = np.zeros(ne_a + 1, dtype=np.int64)
    out = _ip()
    L = lib()
    nnz = L.orc_grid_finder(pos_a.shape[1], int(closed), nthreads, ne_a, enl_a.shape[1], _d(pos_a), _i(enl_a),
                            enl_b.shape[0], enl_b.shape[1], _d(pos_b), _i(enl_b), indptr.ctypes.data_as(_lp),
                            ctypes.byref(out))
    indices = np.ctypeslib.as_array(out, shape=(max(nnz, 1),))[:nnz].copy()
    L.orc_free(ctypes.cast(out, ctypes.c_void_p))
    return indptr, indices


def tree_finder(pos_a, enl_a, pos_b, enl_b):
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    dim = pos_a.shape[1]
    ne_a, ne_b = enl_a.shape[0], enl_b.shape[0]
    indptr = np.zeros(ne_a + 1, dtype=np.int64)
    L = lib()
    args = (dim, pos_b.shape[0], ne_a, enl_a.shape[1], _d(pos_a), _i(enl_a),
            ne_b, enl_b.shape[1], _d(pos_b), _i(enl_b))
    L.orc_tree_finder(*args, indptr.ctypes.data_as(_lp), None)
    indices = np.zeros(max(int(indptr[-1]), 1), dtype=np.int32)
    L.orc_tree_finder(*args, indptr.ctypes.data_as(_lp), _i(indices))
    return indptr, indices[:indptr[-1]]


def csr_to_pairs(indptr, indices):
    rows = np.repeat(np.arange(len(indptr) - 1, dtype=np.int32), np.diff(indptr))
    return rows, np.asarray(indices, dtype=np.int32)


def clip_pairs(pos_a, enl_a, pos_b, enl_b, pa, pb, cap=0, nthreads=8):
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    pa, pb = _i32(pa), _i32(pb)
    dim = pos_a.shape[1]
    n = pa.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    meas = np.zeros(n, dtype=np.float64)
    esz = 6 if dim == 2 else 12
    elems = np.zeros(n * cap * esz if cap else 1, dtype=np.float64)
    lib().orc_clip_pairs(dim, n, _i(pa), _i(pb), _d(pos_a), _i(enl_a), _d(pos_b), _i(enl_b),
                         _i(counts), _d(meas), cap, _d(elems) if cap else None, nthreads)
    if cap:
        shape = (n, cap, 3, 2) if dim == 2 else (n, cap, 4, 3)
        return counts, meas, elems.reshape(shape)
    return counts, meas


def mass_pairs(pos_a, enl_a, pos_b, enl_b, pa, pb, nthreads=8):
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    pa, pb = _i32(pa), _i32(pb)
    dim = pos_a.shape[1]
    loc = dim + 1
    blocks = np.zeros((pa.shape[0], loc, loc), dtype=np.float64)
    lib().orc_mass_pairs(dim, pa.shape[0], _i(pa), _i(pb), _d(pos_a), _i(enl_a), _d(pos_b),
                         _i(enl_b), _d(blocks), nthreads)
    return blocks


def inner_product(pos_a, enl_a, field_a, pos_b, enl_b, field_b, pa, pb):
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    fa, fb = _f64(field_a), _f64(field_b)
    pa, pb = _i32(pa), _i32(pb)
    m = np.zeros(1)
    s = np.zeros(1)
    lib().orc_inner_product(pos_a.shape[1], pa.shape[0], _i(pa), _i(pb), _d(pos_a), _i(enl_a),
                            _d(fa), _d(pos_b), _i(enl_b), _d(fb), _d(m), _d(s))
    return float(m[0]), float(s[0])


def supermesh(pos_a, enl_a, pos_b, enl_b, finder=0, nthreads=1, mass=False):
    """finder 0: tree (strict); 1: brute force strict; 2: brute force closed.
    Returns (stats dict, sum, (build_s, query_clip_s)); with mass the stats
    also hold "fmg", f^T M g for f = x on A and g = y on B."""
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i32(enl_a), _i32(enl_b)
    stats = np.zeros(4, dtype=np.int64)
    s = np.zeros(2)
    t = np.zeros(2)
    lib().orc_supermesh(pos_a.shape[1], finder, nthreads, int(mass), pos_a.shape[0], enl_a.shape[0],
                        _d(pos_a), _i(enl_a), pos_b.shape[0], enl_b.shape[0], _d(pos_b),
                        _i(enl_b), stats.ctypes.data_as(_llp), _d(s), _d(t))
    keys = ("candidates", "nonempty", "elements", "max_per_pair")
    out = dict(zip(keys, (int(x) for x in stats)))
    if mass:
        out["fmg"] = float(s[1])
    return out, float(s[0]), (float(t[0]), float(t[1]))


def load_fixture_meshes():
    z = np.load(os.path.join(ROOT, "tests", "golden", "meshes.npz"))
    names = sorted({k.split("/")[0] for k in z.files})
    return {n: (z[n + "/pos"], z[n + "/enl"]) for n in names}


# -- mesh adjacency and the advancing front finder -------------------------

def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _lpp(a):
    return a.ctypes.data_as(_lp)


def mesh_nelist(nnodes, enl):
    enl = _i64(enl)
    ne, loc = enl.shape
    indptr = np.zeros(nnodes + 1, dtype=np.int64)
    indices = np.zeros(max(ne * loc, 1), dtype=np.int64)
    lib().orc_mesh_nelist(nnodes, ne, loc, _lpp(enl), _lpp(indptr), _lpp(indices))
    return indptr, indices[:ne * loc]


def mesh_nnlist(nnodes, enl):
    enl = _i64(enl)
    ne, loc = enl.shape
    indptr = np.zeros(nnodes + 1, dtype=np.int64)
    indices = np.zeros(max(ne * loc * (loc - 1), 1), dtype=np.int64)
    m = lib().orc_mesh_nnlist(nnodes, ne, loc, _lpp(enl), _lpp(indptr), _lpp(indices))
    return indptr, indices[:m]


def mesh_eelist(nnodes, enl, sloc, nneigh):
    enl = _i64(enl)
    ne, loc = enl.shape
    indptr = np.zeros(ne + 1, dtype=np.int64)
    indices = np.zeros(max(ne * nneigh, 1), dtype=np.int64)
    m = lib().orc_mesh_eelist(nnodes, ne, loc, sloc, nneigh, _lpp(enl), _lpp(indptr), _lpp(indices))
    return indptr, indices[:m]


def transpose_sparsity(ncols, indptr, indices):
    indptr, indices = _i64(indptr), _i64(indices)
    nrows = len(indptr) - 1
    ptr_t = np.zeros(ncols + 1, dtype=np.int64)
    idx_t = np.zeros(max(len(indices), 1), dtype=np.int64)
    lib().orc_transpose_sparsity(nrows, ncols, _lpp(indices), _lpp(indptr), _lpp(idx_t), _lpp(ptr_t))
    return ptr_t, idx_t[:ptr_t[-1]]


def advancing_front(pos_a, enl_a, pos_b, enl_b):
    pos_a, pos_b = _f64(pos_a), _f64(pos_b)
    enl_a, enl_b = _i64(enl_a), _i64(enl_b)
    dim = pos_a.shape[1]
    indptr = np.zeros(enl_a.shape[0] + 1, dtype=np.int64)
    out = _lp()
    L = lib()
    n = L.orc_advancing_front_finder(dim, pos_a.shape[0], enl_a.shape[0], enl_a.shape[1], _d(pos_a),
                                     _lpp(enl_a), pos_b.shape[0], enl_b.shape[0], enl_b.shape[1],
                                     _d(pos_b), _lpp(enl_b), _lpp(indptr), ctypes.byref(out))
    indices = np.ctypeslib.as_array(out, shape=(max(n, 1),))[:n].copy()
    L.orc_free(ctypes.cast(out, ctypes.c_void_p))
    return indptr, indices
