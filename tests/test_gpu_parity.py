"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

Bars: candidate pair sets (sorted rows) bit-exact; per-pair output counts
bit-exact; per-pair measures and output element coordinates bit-exact (the
clippers round exactly like the reference); global sums within 1e-12; mixed
mass within 1e-12 relative.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from libsupermesh_amd import gpu, meshgen

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(np.float64).eps


def fequals(a, b):
    tol = 100.0 * max(EPS, O.spacing(a), O.spacing(b))
    return abs(a - b) < tol


@pytest.fixture(scope="module")
def ctx():
    c = gpu.Context(0)
    yield c
    c.close()


def run(ctx, pa, ea, pb, eb, finder=gpu.FINDER_TREE, flags=0):
    A = gpu.DeviceMesh(pa, ea)
    B = gpu.DeviceMesh(pb, eb)
    if flags & gpu.MASS:
        flags |= gpu.BLOCKS
    r = ctx.supermesh(A, B, finder=finder, flags=flags | gpu.SORTED)
    dim = pa.shape[1]
    what = ["indptr", "indices", "rows", "counts", "measure"]
    if flags & gpu.ELEMENTS:
        what += ["offsets", "elements"]
    if flags & gpu.MASS:
        what += ["blocks"]
    out = gpu.fetch(r, dim, what)
    return r, out


def sorted_csr(indptr, indices):
    idx = np.array(indices, copy=True)
    for a in range(len(indptr) - 1):
        idx[indptr[a]:indptr[a + 1]].sort()
    return idx


def check_against_oracle(pa, ea, pb, eb, r, out, closed=False, elements=False):
    ip, ix = O.brute_force(pa, ea, pb, eb, closed=closed)
    assert np.array_equal(out["indptr"], ip)
    assert np.array_equal(out["indices"], ix)   # brute force lists ascending B ids
    rows, cols = O.csr_to_pairs(ip, ix)
    assert np.array_equal(out["rows"], rows)
    cnt, meas = O.clip_pairs(pa, ea, pb, eb, rows, cols)
    assert np.array_equal(out["counts"], cnt)
    assert np.array_equal(out["measure"], meas)   # bitwise
    assert r.nonempty == int((cnt > 0).sum())
    assert r.nout == int(cnt.sum())
    assert r.max_count == int(cnt.max(initial=0))
    if elements:
        offs = out["offsets"]
        assert offs[-1] == r.nout
        assert np.array_equal(offs[1:], np.cumsum(cnt))
        cap = max(int(cnt.max(initial=0)), 1)
        nz = np.nonzero(cnt)[0]
        for c0 in range(0, len(nz), 20000):   # bounded memory
            sel = nz[c0:c0 + 20000]
            _, _, els = O.clip_pairs(pa, ea, pb, eb, rows[sel], cols[sel], cap=cap)
            for k, p in enumerate(sel):
                got = out["elements"][offs[p]:offs[p + 1]]
                assert np.array_equal(got, els[k, :cnt[p]]), p  # bitwise, reference order
    return cnt, meas


# --------------------------------------------------------------------------
# The reference's fixture meshes and known answers
# --------------------------------------------------------------------------

def test_reference_tri_fixture(ctx, fixture_meshes):
    ref = json.load(open(os.path.join(HERE, "golden", "reference_stats.json")))["triangle_0_01 x square_0_01"]
    pa, ea = fixture_meshes[ref["a"]]
    pb, eb = fixture_meshes[ref["b"]]
    r, out = run(ctx, pa, ea, pb, eb, finder=gpu.FINDER_RTREE)
    assert r.npairs == ref["candidates_closed"]
    assert r.nonempty == ref["nonempty"]
    assert r.nout == ref["elements"]
    assert r.max_count == ref["max_per_pair"]
    assert fequals(r.sum, 0.5)
    check_against_oracle(pa, ea, pb, eb, r, out, closed=True)


def test_reference_tet_fixture(ctx, fixture_meshes):
    ref = json.load(open(os.path.join(HERE, "golden", "reference_stats.json")))["cube_0_05 x pyramid_0_05"]
    pa, ea = fixture_meshes[ref["a"]]
    pb, eb = fixture_meshes[ref["b"]]
    r, out = run(ctx, pa, ea, pb, eb, finder=gpu.FINDER_RTREE, flags=gpu.ELEMENTS)
    assert r.npairs == ref["candidates_closed"]
    assert r.nonempty == ref["nonempty"]
    assert r.nout == ref["elements"]
    assert r.max_count == ref["max_per_pair"]
    assert fequals(r.sum, 1.0 / 3.0)
    check_against_oracle(pa, ea, pb, eb, r, out, closed=True, elements=True)


def test_reference_tiny_meshes(ctx, fixture_meshes):
    for a, b, s in [("cube.1", "cube.2", 1.0), ("square.1", "square.2", 1.0),
                    ("triangle.1", "triangle.2", None)]:
        pa, ea = fixture_meshes[a]
        pb, eb = fixture_meshes[b]
        r, out = run(ctx, pa, ea, pb, eb, flags=gpu.ELEMENTS)
        check_against_oracle(pa, ea, pb, eb, r, out, elements=True)
        if s is not None:
            assert fequals(r.sum, s)


# --------------------------------------------------------------------------
# Synthetic split grids
# --------------------------------------------------------------------------

def test_tri_64_vs_50(ctx):
    pa, ea, pb, eb = meshgen.config_meshes("tri_64_vs_50")
    r, out = run(ctx, pa, ea, pb, eb, flags=gpu.ELEMENTS)
    check_against_oracle(pa, ea, pb, eb, r, out, elements=True)
    assert abs(r.sum - 1.0) <= 1e-12


@pytest.mark.parametrize("na,nb", [(10, 9), (7, 13)])
def test_tet_grids(ctx, na, nb):
    pa, ea = meshgen.split_grid(3, na, 20 + na)
    pb, eb = meshgen.split_grid(3, nb, 40 + nb, mesh_id=1)
    r, out = run(ctx, pa, ea, pb, eb, flags=gpu.ELEMENTS)
    check_against_oracle(pa, ea, pb, eb, r, out, elements=True)
    assert abs(r.sum - 1.0) <= 1e-12


def test_closed_finder_on_touching_grids(ctx):
    # unjittered grids: boxes touch exactly, strict and closed sets differ
    pa, ea = meshgen.split_grid(2, 16, 0, alpha=0.0)
    pb, eb = meshgen.split_grid(2, 8, 0, alpha=0.0, diagonal=1)
    for finder, closed in [(gpu.FINDER_TREE, False), (gpu.FINDER_RTREE, True)]:
        r, out = run(ctx, pa, ea, pb, eb, finder=finder, flags=gpu.ELEMENTS)
        check_against_oracle(pa, ea, pb, eb, r, out, closed=closed, elements=True)
        assert abs(r.sum - 1.0) <= 1e-12


@pytest.mark.parametrize("dim,na,nb,alpha", [(2, 6, 40, None), (2, 4, 32, 0.0), (3, 3, 12, None), (3, 2, 10, 0.0)])
def test_coarse_a_over_fine_b(ctx, dim, na, nb, alpha):
    # A with under an eighth of B's elements: B's elements query a hierarchy
    # over A and the pairs are regrouped by A.  Same pair sets, rows, per-pair
    # results and elements as the oracle, with the strict and the closed
    # predicate (unjittered grids: boxes that touch exactly), and the rows
    # come out sorted without SORTED too.
    pa, ea = meshgen.split_grid(dim, na, 7 + na, alpha=alpha)
    pb, eb = meshgen.split_grid(dim, nb, 11 + nb, alpha=alpha, mesh_id=1)
    assert 8 * ea.shape[0] < eb.shape[0]
    for finder, closed in [(gpu.FINDER_TREE, False), (gpu.FINDER_RTREE, True)]:
        r, out = run(ctx, pa, ea, pb, eb, finder=finder, flags=gpu.ELEMENTS)
        check_against_oracle(pa, ea, pb, eb, r, out, closed=closed, elements=True)
        assert abs(r.sum - 1.0) <= 1e-12
        r2 = ctx.supermesh(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb), finder=finder)
        out2 = gpu.fetch(r2, dim, ["indptr", "indices", "rows", "counts", "measure"])
        for k in out2:
            assert np.array_equal(out2[k], out[k]), k


def test_default_row_order_same_pairs_and_results(ctx):
    # Without SORTED, rows keep the traversal order: same pair set and the
    # same per-pair results once both sides are ordered by (row, col).
    pa, ea = meshgen.split_grid(3, 9, 5)
    pb, eb = meshgen.split_grid(3, 8, 6, mesh_id=1)
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    r = ctx.supermesh(A, B, flags=gpu.ELEMENTS)
    out = gpu.fetch(r, 3, ["indptr", "indices", "rows", "counts", "measure", "offsets", "elements"])
    ip, ix = O.brute_force(pa, ea, pb, eb)
    assert np.array_equal(out["indptr"], ip)
    assert np.array_equal(sorted_csr(out["indptr"], out["indices"]), ix)
    order = np.lexsort((out["indices"], out["rows"]))
    rows, cols = O.csr_to_pairs(ip, ix)
    cnt, meas = O.clip_pairs(pa, ea, pb, eb, rows, cols)
    assert np.array_equal(out["counts"][order], cnt)
    assert np.array_equal(out["measure"][order], meas)
    offs = out["offsets"]
    _, _, els = O.clip_pairs(pa, ea, pb, eb, rows, cols, cap=max(int(cnt.max()), 1))
    for k, p in enumerate(order):
        assert np.array_equal(out["elements"][offs[p]:offs[p + 1]], els[k, :cnt[k]])


def test_deterministic_sum(ctx):
    pa, ea = meshgen.split_grid(3, 12, 3)
    pb, eb = meshgen.split_grid(3, 11, 4, mesh_id=1)
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    s1 = ctx.supermesh(A, B).sum
    s2 = ctx.supermesh(A, B).sum
    assert s1 == s2


# --------------------------------------------------------------------------
# Mixed mass
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dim", [2, 3])
def test_mass_blocks_and_csr(ctx, dim):
    import scipy.sparse as sp
    if dim == 2:
        pa, ea, pb, eb = meshgen.config_meshes("tri_64_vs_50")
    else:
        pa, ea = meshgen.split_grid(3, 8, 11)
        pb, eb = meshgen.split_grid(3, 7, 12, mesh_id=1)
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    # MASS alone: the assembled matrix (2D: no per-pair blocks kept)
    r0 = ctx.supermesh(A, B, flags=gpu.MASS)
    assert (r0.blocks is None) == (dim == 2)
    csr0 = gpu.fetch_mass(r0, pa.shape[0])
    r = ctx.supermesh(A, B, flags=gpu.MASS | gpu.BLOCKS)
    out = gpu.fetch(r, dim, ["rows", "indices", "counts", "blocks"])
    want = O.mass_pairs(pa, ea, pb, eb, out["rows"], out["indices"])
    scale = np.abs(want).max()
    assert np.all(np.abs(out["blocks"] - want) <= 1e-12 * scale)
    indptr, indices, values = gpu.fetch_mass(r, pa.shape[0])
    for x, y in zip(csr0, (indptr, indices, values)):   # the same matrix either way
        assert np.array_equal(x, y)
    M = sp.csr_matrix((values, indices, indptr), shape=(pa.shape[0], pb.shape[0]))
    loc = dim + 1
    ne = out["counts"] > 0
    rr = np.repeat(ea[out["rows"][ne]], loc, axis=1).ravel()
    cc = np.tile(eb[out["indices"][ne]], (1, loc)).ravel()
    W = sp.coo_matrix((want[ne].ravel(), (rr, cc)), shape=M.shape).tocsr()
    W.sum_duplicates()
    assert np.array_equal(np.sort(M.indices), np.sort(W.indices)) or M.nnz == W.nnz
    assert abs(M - W).max() <= 1e-12 * scale
    assert abs(M.sum() - 1.0) < 1e-12
    f, g = pa[:, 0], pb[:, 1]
    assert abs(f @ (M @ g) - 0.25) < 1e-12


# Mesh-size ratios give mass rows of chosen lengths, so that every row kernel
# runs.  2D: node i's row holds 3 slots per candidate pair of each element
# around i (up to 320 slots, up to 512, gathered rows up to 2048, the global
# sort beyond).  3D: 4 entries per non-empty pair (up to 512, up to 2048, the
# global sort).  Rows with more than 256 columns also take the global sort.
@pytest.mark.parametrize("dim,na,nb,lo,hi", [(2, 12, 20, 321, 512), (2, 8, 24, 513, 2048),
                                             (2, 4, 40, 2049, 10 ** 9), (3, 1, 2, 1, 512),
                                             (3, 2, 3, 513, 2048), (3, 6, 8, 2049, 10 ** 9)])
def test_mass_long_rows(ctx, dim, na, nb, lo, hi):
    import scipy.sparse as sp
    pa, ea = meshgen.split_grid(dim, na, 5)
    pb, eb = meshgen.split_grid(dim, nb, 6, mesh_id=1)
    loc = dim + 1
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    r = ctx.supermesh(A, B, flags=gpu.MASS)
    out = gpu.fetch(r, dim, ["indptr", "rows", "indices", "counts"])
    if dim == 2:
        per_element = np.diff(out["indptr"])
    else:
        per_element = np.bincount(out["rows"][out["counts"] > 0], minlength=ea.shape[0])
    raw = np.zeros(pa.shape[0], np.int64)
    for v in range(loc):
        np.add.at(raw, ea[:, v], loc * per_element)
    assert lo <= raw.max() <= hi, raw.max()
    want = O.mass_pairs(pa, ea, pb, eb, out["rows"], out["indices"])
    scale = np.abs(want).max()
    indptr, indices, values = gpu.fetch_mass(r, pa.shape[0])
    assert np.all(np.diff(indptr) >= 0)
    for i in range(pa.shape[0]):   # columns ascending within each row
        assert np.all(np.diff(indices[indptr[i]:indptr[i + 1]]) > 0)
    M = sp.csr_matrix((values, indices, indptr), shape=(pa.shape[0], pb.shape[0]))
    ne = out["counts"] > 0
    rr = np.repeat(ea[out["rows"][ne]], loc, axis=1).ravel()
    cc = np.tile(eb[out["indices"][ne]], (1, loc)).ravel()
    W = sp.coo_matrix((want[ne].ravel(), (rr, cc)), shape=M.shape).tocsr()
    W.sum_duplicates()
    assert M.nnz == W.nnz
    assert abs(M - W).max() <= 1e-12 * scale
    assert abs(M.sum() - 1.0) < 1e-12
    assert abs(pa[:, 0] @ (M @ pb[:, 1]) - 0.25) < 1e-12


def test_reference_integrals_gpu(ctx, fixture_meshes):
    ref = json.load(open(os.path.join(HERE, "golden", "reference_stats.json")))
    for key, fa_of, dim in [("integral_2d", lambda p: p[:, 0], 2),
                            ("integral_3d", lambda p: p[:, 0] + p[:, 2], 3)]:
        r2 = ref[key]
        pa, ea = fixture_meshes[r2["a"]]
        pb, eb = fixture_meshes[r2["b"]]
        A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
        r = ctx.supermesh(A, B, finder=gpu.FINDER_RTREE, flags=gpu.MASS)
        indptr, indices, values = gpu.fetch_mass(r, pa.shape[0])
        import scipy.sparse as sp
        M = sp.csr_matrix((values, indices, indptr), shape=(pa.shape[0], pb.shape[0]))
        val = fa_of(pa) @ (M @ pb[:, 1])
        assert fequals(val, r2["value"]), (key, val)


# --------------------------------------------------------------------------
# Drop-in per-pair and finder functions
# --------------------------------------------------------------------------

def test_dropin_tree_finder(fixture_meshes):
    pa, ea = fixture_meshes["triangle_0_01"]
    pb, eb = fixture_meshes["square_0_01"]
    ip, ix = gpu.tree_intersection_finder(pa, ea, pb, eb)
    rip, rix = O.brute_force(pa, ea, pb, eb)
    assert np.array_equal(ip, rip)
    assert np.array_equal(ix, rix)
    ip, ix = gpu.rtree_intersection_finder(pa, ea, pb, eb)
    rip, rix = O.brute_force(pa, ea, pb, eb, closed=True)
    assert np.array_equal(ip, rip)
    assert np.array_equal(ix, rix)


def test_dropin_per_pair_functions():
    rng = np.random.default_rng(11)
    for _ in range(60):
        ta, tb = rng.uniform(0, 1, (3, 2)), rng.uniform(0, 1, (3, 2))
        assert np.array_equal(gpu.intersect_tris_real(ta, tb), O.intersect_tris(ta, tb))
        assert gpu.triangle_area(ta) == O.triangle_area(ta)
    for _ in range(40):
        ta, tb = rng.uniform(0, 1, (4, 3)), rng.uniform(0, 1, (4, 3))
        assert np.array_equal(gpu.intersect_tets_real(ta, tb), O.intersect_tets(ta, tb))
        assert gpu.tetrahedron_volume(ta) == O.tet_volume_real(ta)


# --------------------------------------------------------------------------
# Edge cases and errors
# --------------------------------------------------------------------------

def test_empty_meshes(ctx):
    pa, ea = meshgen.split_grid(2, 4, 0)
    empty = np.zeros((0, 3), dtype=np.int32)
    r, out = run(ctx, pa, empty, pa, ea)
    assert r.npairs == 0 and r.nelements_a == 0
    r, out = run(ctx, pa, ea, pa, empty)
    assert r.npairs == 0 and np.all(out["indptr"] == 0)


def test_disjoint_meshes(ctx):
    pa, ea = meshgen.split_grid(2, 4, 0)
    r, out = run(ctx, pa, ea, pa + 5.0, ea)
    assert r.npairs == 0 and r.sum == 0.0


def test_invalid_inputs_raise(ctx):
    pa, ea = meshgen.split_grid(2, 4, 0)
    bad = ea.copy()
    bad[3, 1] = pa.shape[0] + 7
    with pytest.raises(gpu.GpuError, match="Invalid node"):
        run(ctx, pa, bad, pa, ea)
    bad_b = ea.copy()
    bad_b[0, 0] = -1
    with pytest.raises(gpu.GpuError, match="Invalid node"):  # in B, with mass
        ctx.supermesh(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pa, bad_b), flags=gpu.MASS)
    p3, e3 = meshgen.split_grid(3, 2, 0)
    with pytest.raises(gpu.GpuError, match="Invalid dimension"):
        run(ctx, pa, ea, p3, e3)
    # the context is still usable afterwards
    r, out = run(ctx, pa, ea, pa, ea)
    assert abs(r.sum - 1.0) < 1e-14


@pytest.mark.parametrize("names", [("triangle.1", "triangle.2"), ("cube.1", "cube.2")])
def test_callback_driver(fixture_meshes, names):
    """Callback driver: B-major order, A candidates ascending (closed overlap),
    every pair's pieces equal to the oracle's intersect of (A, B)."""
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    ctx = gpu.Context()
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    seen = []

    def fn(xa, xb, xc, nodes_b, ia, ib):
        assert np.array_equal(xa, pa[ea[ia]]) and np.array_equal(xb, pb[eb[ib]])
        assert np.array_equal(nodes_b, eb[ib])
        ref = O.intersect_tris(xa, xb) if pa.shape[1] == 2 else O.intersect_tets(xa, xb)
        assert np.array_equal(xc, ref)
        seen.append((ib, ia))

    ctx.supermesh_callback(A, B, fn)
    ip, ix = O.brute_force(pb, eb, pa, ea, closed=True)   # B rows, A ids ascending
    rows, cols = O.csr_to_pairs(ip, ix)
    expect = [(int(r), int(c)) for r, c in zip(rows, cols)]
    assert seen == expect


@pytest.mark.parametrize("names", [("square.1", "square.4"), ("square.3", "square.2"),
                                   ("cube.1", "cube.3"), ("cube.3", "cube.4")])
def test_callback_driver_mixed_elements(fixture_meshes, names):
    """Callback driver on quadrilateral / hexahedral meshes: the whole elements
    reach the callback and every pair's pieces equal the oracle's
    intersect_elements dispatch (Supermesh.F90:162-209)."""
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    dim = pa.shape[1]
    ctx = gpu.Context()
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    seen = []
    total = [0.0]

    def ref_pieces(xa, xb):
        la, lb = len(xa), len(xb)
        if dim == 2:
            if la == 3 and lb == 3:
                return O.intersect_tris(xa, xb)
            if la == 3:
                return O.intersect_tri_quad(xa, xb)
            if lb == 3:
                return O.intersect_tri_quad(xb, xa)
            return O.intersect_quads(xa, xb)
        if la == 4 and lb == 4:
            return O.intersect_tets(xa, xb)
        if la == 4:
            return O.intersect_tet_hex(xa, xb)
        if lb == 4:
            return O.intersect_tet_hex(xb, xa)
        return O.intersect_hexes(xa, xb)

    def fn(xa, xb, xc, nodes_b, ia, ib):
        assert np.array_equal(xa, pa[ea[ia]]) and np.array_equal(xb, pb[eb[ib]])
        assert np.array_equal(nodes_b, eb[ib])
        assert np.array_equal(xc, ref_pieces(xa, xb))
        for t in xc:
            total[0] += O.triangle_area(t) if dim == 2 else O.tet_volume_real(t)
        seen.append((ib, ia))

    ctx.supermesh_callback(A, B, fn)
    ip, ix = O.brute_force(pb, eb, pa, ea, closed=True)
    rows, cols = O.csr_to_pairs(ip, ix)
    assert seen == [(int(r), int(c)) for r, c in zip(rows, cols)]
    expect = {"square": 1.0, "cube": 0.5}[names[0].split(".")[0]]
    assert abs(total[0] - expect) < 1e-12


def test_clip_pairs_rejects_bad_pair_lists():
    """Explicit pair lists: ids outside the meshes and lists of 2^31 pairs or
    more are refused before any kernel reads through them."""
    pa, ea = meshgen.split_grid(2, 4, 0)
    ctx = gpu.Context()
    A = gpu.DeviceMesh(pa, ea)
    good = gpu.DeviceArray(np.zeros(3, dtype=np.int32))
    bad = gpu.DeviceArray(np.array([0, 1, ea.shape[0]], dtype=np.int32))
    with pytest.raises(gpu.GpuError, match="out of range"):
        ctx.clip_pairs(A, A, good, bad)
    with pytest.raises(gpu.GpuError, match="out of range"):
        ctx.clip_pairs(A, A, gpu.DeviceArray(np.array([0, -1, 2], dtype=np.int32)), good)
    r = gpu.Result()
    rc = gpu.lib().libsupermesh_gpu_clip_pairs(ctx.handle, ctypes.byref(A.struct), ctypes.byref(A.struct),
                                               1 << 31, good.ptr, good.ptr, 0, ctypes.byref(r))
    assert rc == 4 and b"2^31" in gpu.lib().libsupermesh_gpu_last_error()
    r = ctx.clip_pairs(A, A, good, good)   # the context is still usable
    assert r.npairs == 3


def test_device_checksum_matches_numpy():
    rng = np.random.default_rng(12)
    host = rng.integers(0, 2**32, 100003, dtype=np.uint64).astype(np.uint32)
    d = gpu.DeviceArray(host)
    ctx = gpu.Context()
    i = np.arange(host.size, dtype=np.uint64)
    ref = int(np.sum(host.astype(np.uint64) * (2 * i + 1), dtype=np.uint64))
    assert ctx.checksum(d.ptr.value, host.nbytes) == ref
    host2 = host.copy()
    host2[[5, 6]] = host2[[6, 5]]   # order matters
    assert ctx.checksum(gpu.DeviceArray(host2).ptr.value, host.nbytes) != ref
