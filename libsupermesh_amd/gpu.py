"""ctypes binding of the C ABI (include/libsupermesh-c.h, include/libsupermesh_gpu.h).

Two layers, both thin:

* the reference's drop-in functions on host arrays -- ``tree_intersection_finder``,
  ``rtree_intersection_finder`` (set_input / query_output / get_output),
  ``intersect_tris_real``, ``intersect_tets_real``, ``triangle_area``,
  ``tetrahedron_volume`` -- with the reference's argument meaning;
* the bulk device API (``Context.supermesh`` / ``Context.intersection_finder`` /
  ``Context.clip_pairs``) on meshes already resident in device memory.

Device memory comes from the library's own HIP runtime
(libsupermesh_gpu_malloc / _memcpy): a second GPU runtime in the same process
(for instance one bundled with another Python package) would compete for the
device.  Every computation runs in the HIP library, and loading fails loudly
if it was not built.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_long, c_void_p

import numpy as np

from . import load

FINDER_TREE = 0
FINDER_RTREE = 1
ELEMENTS = 1
MASS = 2
TIMING = 4
SORTED = 8
LEARNED = 16
BLOCKS = 32
DENSE = 64
GRAPH = 128
PREDICTOR_PARAMS = 12065

_lib = None


class GpuError(RuntimeError):
    """A failure reported by the library (code + text of last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"libsupermesh_gpu error {code}: {msg}")
        self.code = code


class Mesh(ctypes.Structure):
    _fields_ = [("dim", c_int), ("nnodes", c_long), ("nelements", c_long),
                ("positions", c_void_p), ("enlist", c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("nelements_a", c_long), ("npairs", c_long), ("indptr", c_void_p),
                ("indices", c_void_p), ("rows", c_void_p), ("counts", c_void_p),
                ("measure", c_void_p), ("offsets", c_void_p), ("elements", c_void_p),
                ("blocks", c_void_p), ("mass_nnz", c_long), ("mass_indptr", c_void_p),
                ("mass_indices", c_void_p), ("mass_values", c_void_p), ("sum", c_double),
                ("nonempty", c_long), ("nout", c_long), ("max_count", c_int),
                ("ms", c_double * 8), ("learned_slots", c_long), ("learned_overflow", c_long),
                ("learned_quantile", c_double), ("elements_size", c_long), ("count_graph", c_int),
                ("dense_route", c_int)]


# libsupermesh_gpu_pair_callback
CALLBACK = ctypes.CFUNCTYPE(None, POINTER(c_double), POINTER(c_double), POINTER(c_double), c_int,
                            POINTER(c_long), c_long, c_long, c_void_p)


def lib():
    global _lib
    if _lib is None:
        L = load()
        L.libsupermesh_gpu_last_error.restype = ctypes.c_char_p
        L.libsupermesh_gpu_last_error_code.restype = c_int
        L.libsupermesh_gpu_context_create.argtypes = [c_int, POINTER(c_void_p)]
        L.libsupermesh_gpu_context_destroy.argtypes = [c_void_p]
        L.libsupermesh_gpu_context_stream.argtypes = [c_void_p]
        L.libsupermesh_gpu_context_stream.restype = c_void_p
        for name in ("libsupermesh_gpu_supermesh", "libsupermesh_gpu_intersection_finder"):
            f = getattr(L, name)
            f.argtypes = [c_void_p, POINTER(Mesh), POINTER(Mesh), c_int, c_int, POINTER(Result)]
            f.restype = c_int
        L.libsupermesh_gpu_supermesh_mixed.argtypes = [c_void_p, POINTER(Mesh), c_int, POINTER(Mesh),
                                                       c_int, c_int, c_int, POINTER(Result)]
        L.libsupermesh_gpu_clip_pairs.argtypes = [c_void_p, POINTER(Mesh), POINTER(Mesh), c_long,
                                                  c_void_p, c_void_p, c_int, POINTER(Result)]
        L.libsupermesh_gpu_clip_pairs.restype = c_int
        L.libsupermesh_gpu_supermesh_callback.argtypes = [c_void_p, POINTER(Mesh), POINTER(Mesh),
                                                          CALLBACK, c_void_p]
        L.libsupermesh_gpu_supermesh_callback.restype = c_int
        L.libsupermesh_gpu_supermesh_callback_mixed.argtypes = [c_void_p, POINTER(Mesh), c_int,
                                                                POINTER(Mesh), c_int, CALLBACK, c_void_p]
        L.libsupermesh_gpu_supermesh_callback_mixed.restype = c_int
        L.libsupermesh_gpu_supermesh_batched.argtypes = [c_void_p, POINTER(MeshBatch),
                                                         POINTER(MeshBatch), c_int, c_int,
                                                         POINTER(Result), POINTER(BatchResult)]
        L.libsupermesh_gpu_supermesh_batched.restype = c_int
        fp = POINTER(ctypes.c_float)
        L.libsupermesh_gpu_predictor_set.argtypes = [c_void_p, fp]
        L.libsupermesh_gpu_predictor_get.argtypes = [c_void_p, fp]
        L.libsupermesh_gpu_predictor_quantile.argtypes = [c_void_p, ctypes.c_float]
        L.libsupermesh_gpu_predictor_collect.argtypes = [c_void_p, POINTER(Mesh), POINTER(Mesh), c_long,
                                                         POINTER(c_long)]
        L.libsupermesh_gpu_predictor_train.argtypes = [c_void_p, c_int, c_int, ctypes.c_float,
                                                       ctypes.c_uint, fp]
        L.libsupermesh_gpu_predictor_clear.argtypes = [c_void_p]
        L.libsupermesh_gpu_predictor_infer.argtypes = [c_void_p, POINTER(Mesh), POINTER(Mesh), c_long,
                                                       c_void_p, c_void_p, c_void_p]
        L.libsupermesh_gpu_predictor_add_samples.argtypes = [c_void_p, c_long, fp, POINTER(c_int)]
        L.libsupermesh_gpu_predictor_arena.argtypes = [c_void_p, ctypes.c_float, c_long]
        L.libsupermesh_gpu_predictor_arena.restype = c_int
        L.libsupermesh_gpu_predictor_spill.argtypes = [c_void_p, ctypes.c_float, c_long]
        L.libsupermesh_gpu_predictor_spill.restype = c_int
        L.libsupermesh_gpu_checksum.argtypes = [c_void_p, c_void_p, ctypes.c_size_t,
                                                POINTER(ctypes.c_ulonglong)]
        L.libsupermesh_gpu_checksum.restype = c_int
        L.libsupermesh_gpu_compact.argtypes = [c_void_p, POINTER(Result)]
        L.libsupermesh_gpu_compact.restype = c_int
        L.libsupermesh_gpu_count_shape.argtypes = [c_void_p, c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                                   POINTER(ctypes.c_float)]
        L.libsupermesh_gpu_count_shape.restype = c_int
        for name in ("set", "get", "quantile", "collect", "train", "clear", "infer", "add_samples"):
            getattr(L, "libsupermesh_gpu_predictor_" + name).restype = c_int
        L.libsupermesh_gpu_malloc.argtypes = [c_int, ctypes.c_size_t, POINTER(c_void_p)]
        L.libsupermesh_gpu_malloc.restype = c_int
        L.libsupermesh_gpu_free.argtypes = [c_void_p]
        L.libsupermesh_gpu_free.restype = c_int
        L.libsupermesh_gpu_memcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_int]
        L.libsupermesh_gpu_memcpy.restype = c_int
        L.libsupermesh_gpu_synchronize.argtypes = [c_int]
        L.libsupermesh_gpu_synchronize.restype = c_int
        _lib = L
    return _lib


def last_error():
    L = lib()
    return L.libsupermesh_gpu_last_error_code(), L.libsupermesh_gpu_last_error().decode()


def _check(rc):
    if rc != 0:
        code, msg = last_error()
        raise GpuError(code, msg)


def _raise_if_error():
    code, msg = last_error()
    if code != 0:
        raise GpuError(code, msg)


def device_to_numpy(ptr, count, dtype):
    """Copy `count` items of `dtype` from device address `ptr` to a new array."""
    out = np.empty(int(count), dtype=dtype)
    if count > 0:
        _check(lib().libsupermesh_gpu_memcpy(out.ctypes.data, ptr, out.nbytes, 2))
    return out


def synchronize(device=0):
    _check(lib().libsupermesh_gpu_synchronize(device))


class DeviceArray:
    """A device allocation holding a copy of a numpy array."""

    def __init__(self, host, device=0):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        self.ptr = c_void_p()
        _check(lib().libsupermesh_gpu_malloc(device, max(host.nbytes, 1), ctypes.byref(self.ptr)))
        if host.nbytes:
            _check(lib().libsupermesh_gpu_memcpy(self.ptr, host.ctypes.data, host.nbytes, 1))

    def numpy(self):
        return device_to_numpy(self.ptr.value, int(np.prod(self.shape)), self.dtype).reshape(self.shape)

    def write(self, host):
        """Overwrite the allocation in place (same device address) with a
        host array of the same shape and type: e.g. a mesh whose nodes moved."""
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.shape == self.shape
        if host.nbytes:
            _check(lib().libsupermesh_gpu_memcpy(self.ptr, host.ctypes.data, host.nbytes, 1))

    def free(self):
        if self.ptr:
            lib().libsupermesh_gpu_free(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Drop-in functions (host arrays, reference semantics)
# ---------------------------------------------------------------------------

def _pd(a):
    return a.ctypes.data_as(POINTER(c_double))


def _pl(a):
    return a.ctypes.data_as(POINTER(c_long))


def _finder(prefix, pos_a, enl_a, pos_b, enl_b):
    L = lib()
    pos_a = np.ascontiguousarray(pos_a, dtype=np.float64)
    pos_b = np.ascontiguousarray(pos_b, dtype=np.float64)
    enl_a = np.ascontiguousarray(enl_a, dtype=np.int64)
    enl_b = np.ascontiguousarray(enl_b, dtype=np.int64)
    nn_a, dim_a = c_long(pos_a.shape[0]), c_int(pos_a.shape[1])
    ne_a, loc_a = c_long(enl_a.shape[0]), c_int(enl_a.shape[1])
    nn_b, dim_b = c_long(pos_b.shape[0]), c_int(pos_b.shape[1])
    ne_b, loc_b = c_long(enl_b.shape[0]), c_int(enl_b.shape[1])
    getattr(L, prefix + "_set_input")(ctypes.byref(nn_a), ctypes.byref(dim_a), ctypes.byref(ne_a),
                                      ctypes.byref(loc_a), ctypes.byref(nn_b), ctypes.byref(dim_b),
                                      ctypes.byref(ne_b), ctypes.byref(loc_b), _pd(pos_a),
                                      _pl(enl_a), _pd(pos_b), _pl(enl_b))
    _raise_if_error()
    nind = c_long(0)
    getattr(L, prefix + "_query_output")(ctypes.byref(nind))
    _raise_if_error()
    indices = np.zeros(max(nind.value, 1), dtype=np.int64)
    indptr = np.zeros(ne_a.value + 1, dtype=np.int64)
    getattr(L, prefix + "_get_output")(ctypes.byref(ne_a), ctypes.byref(nind), _pl(indices), _pl(indptr))
    _raise_if_error()
    return indptr, indices[:nind.value]


def tree_intersection_finder(pos_a, enl_a, pos_b, enl_b):
    """Strict-overlap candidate pairs as CSR ``(indptr, indices)``, 0-based."""
    return _finder("libsupermesh_tree_intersection_finder", pos_a, enl_a, pos_b, enl_b)


def rtree_intersection_finder(pos_a, enl_a, pos_b, enl_b):
    """Closed-overlap (R-tree semantics) candidate pairs as CSR."""
    return _finder("libsupermesh_rtree_intersection_finder", pos_a, enl_a, pos_b, enl_b)


def sort_intersection_finder(pos_a, enl_a, pos_b, enl_b):
    """1D (intervals) candidate pairs of the reference's sort finder, as CSR.
    Host code in the library: no GPU is touched."""
    return _finder("libsupermesh_sort_intersection_finder", pos_a, enl_a, pos_b, enl_b)


class DeviceTree:
    """A search hierarchy over mesh B kept on the GPU, queried one element at
    a time (the reference's tree_type / rtree_type); rows ascending, 0-based."""

    def __init__(self, positions, enlist, finder=FINDER_TREE):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        enl = np.ascontiguousarray(enlist, dtype=np.int64)
        self.dim = pos.shape[1]
        self.handle = c_void_p()
        L = lib()
        L.libsupermesh_gpu_tree_create.argtypes = [c_int, c_int, c_long, c_void_p, c_long, c_int, c_void_p,
                                                   POINTER(c_void_p)]
        L.libsupermesh_gpu_tree_query.argtypes = [c_void_p, c_int, c_void_p, POINTER(c_long)]
        L.libsupermesh_gpu_tree_result.argtypes = [c_void_p, c_long, c_void_p]
        L.libsupermesh_gpu_tree_destroy.argtypes = [c_void_p]
        _check(L.libsupermesh_gpu_tree_create(finder, self.dim, pos.shape[0], pos.ctypes.data, enl.shape[0],
                                              enl.shape[1], enl.ctypes.data, ctypes.byref(self.handle)))

    def query(self, element):
        el = np.ascontiguousarray(element, dtype=np.float64)
        n = c_long(0)
        _check(lib().libsupermesh_gpu_tree_query(self.handle, el.shape[0], el.ctypes.data, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.int64)
        _check(lib().libsupermesh_gpu_tree_result(self.handle, n.value, out.ctypes.data))
        return out[:n.value]

    def close(self):
        if self.handle:
            lib().libsupermesh_gpu_tree_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def advancing_front_intersection_finder(pos_a, enl_a, pos_b, enl_b):
    """Candidate pairs of the reference's advancing front finder (closed
    overlap, rows in the front's visiting order), as CSR.  The front runs on
    the host; the face-neighbour lists come from the GPU."""
    return _finder("libsupermesh_advancing_front_intersection_finder", pos_a, enl_a, pos_b, enl_b)


def _graph_call(prefix, nrows_out, *args):
    L = lib()
    getattr(L, prefix + "_set_input")(*args)
    _raise_if_error()
    nind = c_long(0)
    getattr(L, prefix + "_query_output")(ctypes.byref(nind))
    _raise_if_error()
    indptr = np.zeros(nrows_out + 1, dtype=np.int64)
    indices = np.zeros(max(nind.value, 1), dtype=np.int64)
    getattr(L, prefix + "_get_output")(ctypes.byref(c_long(nrows_out)), ctypes.byref(nind), _pl(indices),
                                       _pl(indptr))
    _raise_if_error()
    return indptr, indices[:nind.value]


def mesh_nelist(nnodes, enlist):
    """Node -> element CSR ``(indptr, indices)`` (mesh_nelist), 0-based."""
    enl = np.ascontiguousarray(enlist, dtype=np.int64)
    ne, loc = enl.shape
    return _graph_call("libsupermesh_mesh_nelist", nnodes, ctypes.byref(c_long(nnodes)),
                       ctypes.byref(c_long(ne)), ctypes.byref(c_long(loc)), _pl(enl))


def mesh_nnlist(nnodes, enlist):
    """Node -> node CSR (nodes sharing an element, ascending), 0-based."""
    enl = np.ascontiguousarray(enlist, dtype=np.int64)
    ne, loc = enl.shape
    return _graph_call("libsupermesh_mesh_nnlist", nnodes, ctypes.byref(c_long(nnodes)),
                       ctypes.byref(c_long(ne)), ctypes.byref(c_long(loc)), _pl(enl))


def mesh_eelist(nnodes, enlist, sloc):
    """Element -> element CSR (elements sharing ``sloc`` nodes, in the
    reference's discovery order, at most loc each), 0-based."""
    enl = np.ascontiguousarray(enlist, dtype=np.int64)
    ne, loc = enl.shape
    return _graph_call("libsupermesh_mesh_eelist", ne, ctypes.byref(c_long(nnodes)),
                       ctypes.byref(c_long(ne)), ctypes.byref(c_long(loc)), ctypes.byref(c_long(sloc)),
                       _pl(enl))


def transpose_sparsity(ncols, indptr, indices):
    """CSR transpose ``(indptr_t, indices_t)``: rows of each column ascending."""
    ptr = np.ascontiguousarray(indptr, dtype=np.int64)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    nrows = len(ptr) - 1
    ptr_t = np.zeros(ncols + 1, dtype=np.int64)
    idx_t = np.zeros(max(len(idx), 1), dtype=np.int64)
    lib().libsupermesh_transpose_sparsity(ctypes.byref(c_long(nrows)), ctypes.byref(c_long(ncols)),
                                          ctypes.byref(c_long(len(idx))), _pl(idx), _pl(ptr), _pl(idx_t),
                                          _pl(ptr_t))
    _raise_if_error()
    return ptr_t, idx_t[:ptr_t[-1]]


def _pair_call(fn, a, b, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(shape)
    n = c_int(0)
    fn(_pd(a), _pd(b), _pd(out), ctypes.byref(n))
    _raise_if_error()
    return out[:n.value].copy()


def intersect_tri_quad(tri_a, quad_b):
    """Triangles of tri_a ^ quad_b, ``float64[n, 3, 2]`` (at most 46)."""
    return _pair_call(lib().libsupermesh_intersect_tri_quad, tri_a, quad_b, (46, 3, 2))


def intersect_tet_hex(tet_a, hex_b):
    """Tetrahedra of tet_a ^ hex_b (Gmsh vertex order), ``float64[n, 4, 3]`` (at most 729)."""
    return _pair_call(lib().libsupermesh_intersect_tet_hex, tet_a, hex_b, (729, 4, 3))


def intersect_quads(quad_a, quad_b):
    """Triangles of quad_a ^ quad_b (intersect_quads), ``float64[n, 3, 2]`` (at most 62)."""
    return _pair_call(lib().libsupermesh_intersect_quads, quad_a, quad_b, (62, 3, 2))


def intersect_hexes(hex_a, hex_b):
    """Tetrahedra of hex_a ^ hex_b (intersect_hexes), ``float64[n, 4, 3]`` (at most 3645)."""
    return _pair_call(lib().libsupermesh_intersect_hexes, hex_a, hex_b, (3645, 4, 3))


def intersect_intervals(interval_a, interval_b):
    """``(n, interval)``: n is 0 or 1.  Host code in the library."""
    a = np.ascontiguousarray(interval_a, dtype=np.float64).reshape(2)
    b = np.ascontiguousarray(interval_b, dtype=np.float64).reshape(2)
    out = np.zeros(2)
    n = c_int(0)
    lib().libsupermesh_intersect_intervals(_pd(a), _pd(b), _pd(out), ctypes.byref(n))
    _raise_if_error()
    return n.value, out


def interval_size(interval):
    v = np.ascontiguousarray(interval, dtype=np.float64).reshape(2)
    r = c_double(0.0)
    lib().libsupermesh_interval_size(_pd(v), ctypes.byref(r))
    _raise_if_error()
    return r.value


def intersect_tris_real(tri_a, tri_b):
    """Triangles of tri_a ^ tri_b, ``float64[n, 3, 2]``."""
    a = np.ascontiguousarray(tri_a, dtype=np.float64)
    b = np.ascontiguousarray(tri_b, dtype=np.float64)
    out = np.zeros((22, 3, 2))
    n = c_int(0)
    lib().libsupermesh_intersect_tris_real(_pd(a), _pd(b), _pd(out), ctypes.byref(n))
    _raise_if_error()
    return out[:n.value].copy()


def intersect_tets_real(tet_a, tet_b):
    """Tetrahedra of tet_a ^ tet_b, ``float64[n, 4, 3]``."""
    a = np.ascontiguousarray(tet_a, dtype=np.float64)
    b = np.ascontiguousarray(tet_b, dtype=np.float64)
    out = np.zeros((81, 4, 3))
    n = c_int(0)
    lib().libsupermesh_intersect_tets_real(_pd(a), _pd(b), _pd(out), ctypes.byref(n))
    _raise_if_error()
    return out[:n.value].copy()


def triangle_area(tri):
    t = np.ascontiguousarray(tri, dtype=np.float64)
    r = c_double(0.0)
    lib().libsupermesh_triangle_area(_pd(t), ctypes.byref(r))
    _raise_if_error()
    return r.value


def tetrahedron_volume(tet):
    t = np.ascontiguousarray(tet, dtype=np.float64)
    r = c_double(0.0)
    lib().libsupermesh_tetrahedron_volume(_pd(t), ctypes.byref(r))
    _raise_if_error()
    return r.value


# ---------------------------------------------------------------------------
# Bulk device API
# ---------------------------------------------------------------------------

class DeviceMesh:
    """A mesh copied to device memory: positions float64[nnodes, dim],
    enlist int32[nelements, loc] (0-based; loc = dim+1 for simplices, 2^dim
    for quadrilaterals / hexahedra)."""

    def __init__(self, positions, enlist, device=0):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        enl = np.ascontiguousarray(enlist, dtype=np.int32)
        self.dim = pos.shape[1]
        self.loc = enl.shape[1]
        self.positions = DeviceArray(pos, device)
        self.enlist = DeviceArray(enl, device)
        self.struct = Mesh(self.dim, pos.shape[0], enl.shape[0], self.positions.ptr.value,
                           self.enlist.ptr.value)

    @property
    def nnodes(self):
        return self.struct.nnodes

    @property
    def nelements(self):
        return self.struct.nelements


class MeshBatch(ctypes.Structure):
    _fields_ = [("dim", c_int), ("nmeshes", c_long), ("nnodes", c_long), ("nelements", c_long),
                ("positions", c_void_p), ("enlist", c_void_p), ("element_offsets", c_void_p)]


class BatchResult(ctypes.Structure):
    _fields_ = [("nmeshes", c_long), ("pair_offsets", c_void_p), ("sums", c_void_p)]


class DeviceMeshBatch:
    """Meshes of one dimension concatenated on the device: node ids are made
    batch-global, ``element_offsets[m]`` is the first element of mesh m."""

    def __init__(self, meshes, device=0):
        dim = meshes[0][0].shape[1]
        node_off = np.zeros(len(meshes) + 1, dtype=np.int64)
        elem_off = np.zeros(len(meshes) + 1, dtype=np.int64)
        for m, (pos, enl) in enumerate(meshes):
            assert pos.shape[1] == dim
            node_off[m + 1] = node_off[m] + pos.shape[0]
            elem_off[m + 1] = elem_off[m] + enl.shape[0]
        pos = np.concatenate([np.asarray(p, dtype=np.float64) for p, _ in meshes])
        enl = np.concatenate([np.asarray(e, dtype=np.int64) + node_off[m]
                              for m, (_, e) in enumerate(meshes)]).astype(np.int32)
        self.dim = dim
        self.node_offsets = node_off
        self.element_offsets = elem_off
        self.positions = DeviceArray(np.ascontiguousarray(pos), device)
        self.enlist = DeviceArray(np.ascontiguousarray(enl), device)
        self.offsets = DeviceArray(elem_off, device)
        self.struct = MeshBatch(dim, len(meshes), pos.shape[0], enl.shape[0],
                                self.positions.ptr.value, self.enlist.ptr.value,
                                self.offsets.ptr.value)


class Context:
    """A library context (its own HIP stream and reusable device buffers).
    Results returned by a call stay valid until the next call."""

    def __init__(self, device=0):
        self.handle = c_void_p()
        _check(lib().libsupermesh_gpu_context_create(device, ctypes.byref(self.handle)))

    def close(self):
        if self.handle:
            lib().libsupermesh_gpu_context_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().libsupermesh_gpu_context_stream(self.handle)

    def supermesh(self, a: DeviceMesh, b: DeviceMesh, finder=FINDER_TREE, flags=0):
        r = Result()
        _check(lib().libsupermesh_gpu_supermesh(self.handle, ctypes.byref(a.struct),
                                                ctypes.byref(b.struct), finder, flags,
                                                ctypes.byref(r)))
        return r

    def supermesh_mixed(self, a: DeviceMesh, b: DeviceMesh, finder=FINDER_RTREE, flags=0):
        """Supermesh of meshes of any element type (triangles / quadrilaterals
        in 2D, tetrahedra / hexahedra in 3D), clipped as intersect_elements."""
        r = Result()
        _check(lib().libsupermesh_gpu_supermesh_mixed(self.handle, ctypes.byref(a.struct), a.loc,
                                                      ctypes.byref(b.struct), b.loc, finder, flags,
                                                      ctypes.byref(r)))
        return r

    def intersection_finder(self, a: DeviceMesh, b: DeviceMesh, finder=FINDER_TREE, flags=0):
        r = Result()
        _check(lib().libsupermesh_gpu_intersection_finder(self.handle, ctypes.byref(a.struct),
                                                          ctypes.byref(b.struct), finder, flags,
                                                          ctypes.byref(r)))
        return r

    def count_shape(self, dim):
        """The count pass's launch shape measured fastest for `dim`
        (GRAPH calls): {block, grid, lds_depths, ms}; zeros before the first."""
        b, g, d = c_int(0), c_int(0), c_int(0)
        ms = ctypes.c_float(0)
        _check(lib().libsupermesh_gpu_count_shape(self.handle, dim, ctypes.byref(b), ctypes.byref(g),
                                                  ctypes.byref(d), ctypes.byref(ms)))
        return {"block": b.value, "grid": g.value, "lds_depths": d.value, "ms": round(ms.value, 4)}

    def compact(self, r: Result):
        """The plain layout (offsets = prefix sums of the counts) of the
        context's latest result, in place; a no-op for a compact one."""
        _check(lib().libsupermesh_gpu_compact(self.handle, ctypes.byref(r)))
        return r

    def checksum(self, ptr, nbytes):
        """Order-sensitive 64-bit checksum of device memory (whole 32-bit words)."""
        h = ctypes.c_ulonglong(0)
        _check(lib().libsupermesh_gpu_checksum(self.handle, c_void_p(ptr), nbytes, ctypes.byref(h)))
        return h.value

    # -- learned output-size predictor -------------------------------------
    def predictor_set(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        assert p.size == PREDICTOR_PARAMS
        _check(lib().libsupermesh_gpu_predictor_set(self.handle, p.ctypes.data_as(POINTER(ctypes.c_float))))

    def predictor_get(self):
        p = np.zeros(PREDICTOR_PARAMS, dtype=np.float32)
        _check(lib().libsupermesh_gpu_predictor_get(self.handle, p.ctypes.data_as(POINTER(ctypes.c_float))))
        return p

    def predictor_quantile(self, q):
        _check(lib().libsupermesh_gpu_predictor_quantile(self.handle, ctypes.c_float(q)))

    def predictor_arena(self, fraction, min_tets):
        _check(lib().libsupermesh_gpu_predictor_arena(self.handle, ctypes.c_float(fraction), min_tets))

    def predictor_spill(self, fraction, min_tets):
        _check(lib().libsupermesh_gpu_predictor_spill(self.handle, ctypes.c_float(fraction), min_tets))

    def predictor_collect(self, a: DeviceMesh, b: DeviceMesh, max_samples):
        n = c_long(0)
        _check(lib().libsupermesh_gpu_predictor_collect(self.handle, ctypes.byref(a.struct),
                                                        ctypes.byref(b.struct), max_samples,
                                                        ctypes.byref(n)))
        return n.value

    def predictor_train(self, epochs, batch=4096, lr=2e-3, seed=0):
        loss = np.zeros(max(epochs, 1), dtype=np.float32)
        _check(lib().libsupermesh_gpu_predictor_train(self.handle, epochs, batch, ctypes.c_float(lr),
                                                      seed, loss.ctypes.data_as(POINTER(ctypes.c_float))))
        return loss[:epochs]

    def predictor_infer(self, a: DeviceMesh, b: DeviceMesh, pair_a, pair_b):
        """Predicted output budget of each pair (host int arrays in, numpy out)."""
        pa = DeviceArray(np.ascontiguousarray(pair_a, dtype=np.int32))
        pb = DeviceArray(np.ascontiguousarray(pair_b, dtype=np.int32))
        n = len(pair_a)
        out = DeviceArray(np.zeros(max(n, 1), dtype=np.int32))
        _check(lib().libsupermesh_gpu_predictor_infer(self.handle, ctypes.byref(a.struct),
                                                      ctypes.byref(b.struct), n, pa.ptr, pb.ptr,
                                                      out.ptr))
        return out.numpy()[:n]

    def predictor_add_samples(self, features, counts):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 24)
        y = np.ascontiguousarray(counts, dtype=np.int32).reshape(-1)
        assert f.shape[0] == y.shape[0]
        _check(lib().libsupermesh_gpu_predictor_add_samples(
            self.handle, f.shape[0], f.ctypes.data_as(POINTER(ctypes.c_float)),
            y.ctypes.data_as(POINTER(c_int))))

    def predictor_clear(self):
        _check(lib().libsupermesh_gpu_predictor_clear(self.handle))

    def supermesh_batched(self, a: DeviceMeshBatch, b: DeviceMeshBatch, finder=FINDER_TREE, flags=0):
        """(Result over all meshes, BatchResult with per-mesh pair offsets and sums)."""
        r = Result()
        br = BatchResult()
        _check(lib().libsupermesh_gpu_supermesh_batched(self.handle, ctypes.byref(a.struct),
                                                        ctypes.byref(b.struct), finder, flags,
                                                        ctypes.byref(r), ctypes.byref(br)))
        return r, br

    def supermesh_callback(self, a: DeviceMesh, b: DeviceMesh, fn):
        """The callback driver: ``fn(element_a, element_b, elements_c, nodes_b,
        ele_a, ele_b)`` for every closed-overlap candidate pair, B element by B
        element, A candidates ascending (0-based ids, numpy arrays valid only
        during the call).  Meshes may hold any element type (a.loc, b.loc):
        element_a is [a.loc, dim], element_b and nodes_b follow b.loc, and the
        clipped elements are simplices, as intersect_elements gives them."""
        dim = a.struct.dim
        loc = dim + 1
        la, lb = a.loc, b.loc
        err = []

        def tramp(pa, pb, pc, n, nodes, ea, eb, user):
            if err:
                return
            try:
                A = np.ctypeslib.as_array(pa, shape=(la, dim))
                B = np.ctypeslib.as_array(pb, shape=(lb, dim))
                Cc = (np.ctypeslib.as_array(pc, shape=(n, loc, dim)) if n > 0
                      else np.zeros((0, loc, dim)))
                nb = np.ctypeslib.as_array(nodes, shape=(lb,))
                fn(A, B, Cc, nb, ea, eb)
            except BaseException as e:  # surfaced after the driver returns
                err.append(e)

        cb = CALLBACK(tramp)
        _check(lib().libsupermesh_gpu_supermesh_callback_mixed(self.handle, ctypes.byref(a.struct), la,
                                                               ctypes.byref(b.struct), lb, cb, None))
        if err:
            raise err[0]

    def clip_pairs(self, a: DeviceMesh, b: DeviceMesh, pair_a: DeviceArray, pair_b: DeviceArray,
                   flags=0):
        """pair_a / pair_b: int32 device arrays of equal length."""
        r = Result()
        n = int(np.prod(pair_a.shape))
        _check(lib().libsupermesh_gpu_clip_pairs(self.handle, ctypes.byref(a.struct),
                                                 ctypes.byref(b.struct), n, pair_a.ptr,
                                                 pair_b.ptr, flags, ctypes.byref(r)))
        return r


def fetch(r: Result, dim, what=("indptr", "indices", "rows", "counts", "measure")):
    """Copy the requested result arrays to host numpy arrays."""
    loc = dim + 1
    out = {}
    for w in what:
        if w == "indptr":
            out[w] = device_to_numpy(r.indptr, r.nelements_a + 1, np.int64)
        elif w == "indices":
            out[w] = device_to_numpy(r.indices, r.npairs, np.int32)
        elif w == "rows":
            out[w] = device_to_numpy(r.rows, r.npairs, np.int32)
        elif w == "counts":
            out[w] = device_to_numpy(r.counts, r.npairs, np.int32)
        elif w == "measure":
            out[w] = device_to_numpy(r.measure, r.npairs, np.float64)
        elif w == "offsets":
            out[w] = device_to_numpy(r.offsets, r.npairs + 1, np.int64)
        elif w == "elements":
            out[w] = device_to_numpy(r.elements, r.elements_size * loc * dim, np.float64).reshape(-1, loc, dim)
        elif w == "blocks":
            out[w] = device_to_numpy(r.blocks, r.npairs * loc * loc, np.float64).reshape(-1, loc, loc)
        else:
            raise KeyError(w)
    return out


def fetch_batch(br: BatchResult):
    """Per-mesh pair offsets and measure sums of a batched call, as numpy."""
    n = br.nmeshes
    return (device_to_numpy(br.pair_offsets, n + 1, np.int64),
            device_to_numpy(br.sums, n, np.float64))


def fetch_mass(r: Result, nnodes_a):
    """The assembled mixed mass matrix as CSR arrays (indptr, indices, values)."""
    indptr = device_to_numpy(r.mass_indptr, nnodes_a + 1, np.int64)
    indices = device_to_numpy(r.mass_indices, r.mass_nnz, np.int32)
    values = device_to_numpy(r.mass_values, r.mass_nnz, np.float64)
    return indptr, indices, values
