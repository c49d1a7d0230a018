"""GPU parity of the mixed-element drop-in functions (intersect_tri_quad,
intersect_tet_hex) against the oracle, bit for bit, through the C ABI."""
import numpy as np
import pytest

import oracle_lib as O
from libsupermesh_amd import gpu

pytestmark = pytest.mark.gpu


def _quad(rng):
    c = rng.uniform(0.2, 0.8, 2)
    r = rng.uniform(0.1, 0.5)
    q = c + r * np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]]) + rng.uniform(-0.15, 0.15, (4, 2)) * r
    return q if rng.uniform() < 0.5 else q[::-1].copy()


def _hex(rng):
    c = rng.uniform(0.2, 0.8, 3)
    r = rng.uniform(0.1, 0.5)
    s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    return c + r * s + rng.uniform(-0.1, 0.1, (8, 3)) * r


def test_tri_quad_bit_exact():
    rng = np.random.default_rng(31)
    nonempty = 0
    for i in range(600):
        tri = rng.uniform(0, 1, (3, 2))
        quad = _quad(rng)
        got = gpu.intersect_tri_quad(tri, quad)
        ref = O.intersect_tri_quad(tri, quad)
        assert got.shape == ref.shape, i
        assert np.array_equal(got, ref), i
        nonempty += len(ref) > 0
    assert nonempty > 200


def test_tet_hex_bit_exact():
    rng = np.random.default_rng(32)
    nonempty = 0
    for i in range(300):
        tet = rng.uniform(0, 1, (4, 3))
        hx = _hex(rng)
        got = gpu.intersect_tet_hex(tet, hx)
        ref = O.intersect_tet_hex(tet, hx)
        assert got.shape == ref.shape, i
        assert np.array_equal(got, ref), i
        nonempty += len(ref) > 0
    assert nonempty > 100


def test_quads_and_hexes_bit_exact():
    # the per-pair intersect_quads / intersect_hexes against the oracle
    rng = np.random.default_rng(33)
    nq = nh = 0
    for i in range(300):
        qa, qb = _quad(rng), _quad(rng)
        got, ref = gpu.intersect_quads(qa, qb), O.intersect_quads(qa, qb)
        assert got.shape == ref.shape and np.array_equal(got, ref), i
        nq += len(ref) > 0
    for i in range(60):
        ha, hb = _hex(rng), _hex(rng)
        got, ref = gpu.intersect_hexes(ha, hb), O.intersect_hexes(ha, hb)
        assert got.shape == ref.shape and np.array_equal(got, ref), i
        nh += len(ref) > 0
    assert nq > 100 and nh > 20


@pytest.mark.parametrize("names,expected", [(("square.1", "square.4"), 1.0),
                                            (("cube.1", "cube.3"), 0.5)])
def test_reference_mixed_meshes(fixture_meshes, names, expected):
    # src/tests/test_element_intersector.F90 totals, pair by pair on the GPU
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    indptr, indices = O.brute_force(pa, ea, pb, eb, closed=True)
    rows, cols = O.csr_to_pairs(indptr, indices)
    fn = gpu.intersect_tri_quad if pa.shape[1] == 2 else gpu.intersect_tet_hex
    ref_fn = O.intersect_tri_quad if pa.shape[1] == 2 else O.intersect_tet_hex
    meas = O.triangle_area if pa.shape[1] == 2 else O.tet_volume_real
    total = 0.0
    for a, b in zip(rows, cols):
        got = fn(pa[ea[a]], pb[eb[b]])
        assert np.array_equal(got, ref_fn(pa[ea[a]], pb[eb[b]]))
        total += sum(meas(t) for t in got)
    assert abs(total - expected) < 1e-12


@pytest.mark.parametrize("names", [("square.1", "square.4"), ("square.3", "square.2"),
                                   ("square.3", "square.4"), ("cube.1", "cube.3"),
                                   ("cube.3", "cube.4")])
@pytest.mark.parametrize("closed", [False, True])
def test_finders_on_quads_and_hexes(fixture_meshes, names, closed):
    # the candidate finders take any element type (boxes only)
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    fn = gpu.rtree_intersection_finder if closed else gpu.tree_intersection_finder
    indptr, indices = fn(pa, ea, pb, eb)
    ip, ix = O.brute_force(pa, ea, pb, eb, closed=closed)
    assert np.array_equal(indptr, ip)
    assert np.array_equal(indices, ix)


def _oracle_elements(ea_xy, eb_xy):
    """intersect_elements (Supermesh.F90:162-209) over the oracle's routines."""
    dim, la, lb = ea_xy.shape[1], ea_xy.shape[0], eb_xy.shape[0]
    if dim == 2:
        if la == 3 and lb == 4:
            return O.intersect_tri_quad(ea_xy, eb_xy)
        if la == 4 and lb == 3:
            return O.intersect_tri_quad(eb_xy, ea_xy)
        return O.intersect_quads(ea_xy, eb_xy)
    if la == 4 and lb == 8:
        return O.intersect_tet_hex(ea_xy, eb_xy)
    if la == 8 and lb == 4:
        return O.intersect_tet_hex(eb_xy, ea_xy)
    return O.intersect_hexes(ea_xy, eb_xy)


def _check_bulk(pa, ea, pb, eb, finder, sort, expected=None):
    dim = pa.shape[1]
    ctx = gpu.Context()
    flags = gpu.ELEMENTS | (gpu.SORTED if sort else 0)
    r = ctx.supermesh_mixed(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb), finder=finder, flags=flags)
    out = gpu.fetch(r, dim, ["indptr", "indices", "rows", "counts", "measure", "offsets", "elements"])
    ip, ix = O.brute_force(pa, ea, pb, eb, closed=finder == gpu.FINDER_RTREE)
    assert np.array_equal(out["indptr"], ip)
    if sort:
        assert np.array_equal(out["indices"], ix)
    else:
        for a in range(len(ip) - 1):
            assert sorted(out["indices"][ip[a]:ip[a + 1]]) == list(ix[ip[a]:ip[a + 1]])
    meas = O.triangle_area if dim == 2 else O.tet_volume_real
    total, nout = 0.0, 0
    for p, (a, b) in enumerate(zip(out["rows"], out["indices"])):
        ref = _oracle_elements(pa[ea[a]], pb[eb[b]])
        assert out["counts"][p] == len(ref), p
        m = 0.0
        for t in ref:
            m = m + meas(t)
        assert out["measure"][p] == m, p
        o0, o1 = out["offsets"][p], out["offsets"][p + 1]
        assert o1 - o0 == len(ref)
        assert np.array_equal(out["elements"][o0:o1], ref.reshape(-1, dim + 1, dim)), p
        total += m
        nout += len(ref)
    assert r.nout == nout and r.npairs == len(ix)
    assert abs(r.sum - total) <= 1e-13 * max(1.0, total)
    if expected is not None:
        assert abs(r.sum - expected) < 1e-12
    return r


@pytest.mark.parametrize("names,expected", [
    (("square.1", "square.4"), 1.0), (("square.4", "square.1"), 1.0), (("square.3", "square.2"), 1.0),
    (("square.3", "square.4"), 1.0), (("cube.1", "cube.3"), 0.5), (("cube.3", "cube.2"), 0.5),
    (("cube.3", "cube.4"), 0.5)])
def test_bulk_mixed_reference_meshes(fixture_meshes, names, expected):
    # the meshes and totals of src/tests/test_element_intersector.F90:136-332,
    # every candidate pair clipped in one call, pair by pair equal to the oracle
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    _check_bulk(pa, ea, pb, eb, gpu.FINDER_RTREE, True, expected)


@pytest.mark.parametrize("dim,na,nb", [(2, 9, 7), (3, 3, 4)])
@pytest.mark.parametrize("kind", ["simplex_x_quad", "quad_x_simplex", "quad_x_quad"])
@pytest.mark.parametrize("finder,sort", [(gpu.FINDER_RTREE, True), (gpu.FINDER_TREE, False)])
def test_bulk_mixed_jittered(dim, na, nb, kind, finder, sort):
    from libsupermesh_amd import meshgen
    qa = meshgen.cell_grid(dim, na, 11)
    qb = meshgen.cell_grid(dim, nb, 12, mesh_id=1)
    sa = meshgen.split_grid(dim, na, 5, mesh_id=0, diagonal=0)
    sb = meshgen.split_grid(dim, nb, 5, mesh_id=1, diagonal=1)
    A, B = {"simplex_x_quad": (sa, qb), "quad_x_simplex": (qa, sb), "quad_x_quad": (qa, qb)}[kind]
    r = _check_bulk(A[0], A[1], B[0], B[1], finder, sort)
    assert abs(r.sum - 1.0) < 1e-10


def test_bulk_mixed_rejects_mass_and_bad_loc():
    from libsupermesh_amd import meshgen
    pa, ea = meshgen.cell_grid(2, 3, 1)
    ctx = gpu.Context()
    m = gpu.DeviceMesh(pa, ea)
    with pytest.raises(gpu.GpuError, match="simplex"):
        ctx.supermesh_mixed(m, m, flags=gpu.MASS)
    bad = gpu.DeviceMesh(pa, np.zeros((2, 5), dtype=np.int32))
    with pytest.raises(gpu.GpuError, match="loc"):
        ctx.supermesh_mixed(m, bad)


def test_bulk_mixed_concave_quad():
    # an arrowhead quad whose vertices alternate sides of B's first line: the
    # first pass emits 6 points, more than a convex quad can (the library's
    # per-lane polygon slots), so the pair is clipped by the streaming path;
    # the result must still be the reference's (intersect_quads)
    pa = np.array([[0.0, 0.0], [2.0, 1.0], [0.0, 2.0], [0.5, 1.0]])
    ea = np.array([[0, 1, 2, 3]], dtype=np.int32)
    pb = np.array([[0.3, -1.0], [0.3, 3.0], [-1.0, 3.0], [-1.0, -1.0]])
    eb = np.array([[0, 1, 2, 3]], dtype=np.int32)
    ref = O.intersect_quads(pa[ea[0]], pb[eb[0]])
    assert len(ref) >= 3
    r = _check_bulk(pa, ea, pb, eb, gpu.FINDER_RTREE, True)
    assert r.nout == len(ref)
