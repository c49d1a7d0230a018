"""Pins of the oracle's mixed element types (tri-quad, quads, tet-hex, hexes)
and 1D routines (intervals, sort finder) against the reference tests' known
answers, exact arithmetic and closed forms.  CPU only."""
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O
from test_oracle import fequals

# Unit cube in Gmsh hexahedron order.
CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)
# The 6 Kuhn tets of the unit cube (Gmsh vertex ids), around the 0-6 diagonal.
KUHN = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]


def _pairs_closed(pa, ea, pb, eb):
    indptr, indices = O.brute_force(pa, ea, pb, eb, closed=True)
    return zip(*O.csr_to_pairs(indptr, indices))


def _sum_mixed(fixture_meshes, name_a, name_b):
    """The loop of test_element_intersector.F90 over intersect_elements for
    one mesh pair: dispatch on loc as Supermesh.F90:162-209 does."""
    pa, ea = fixture_meshes[name_a]
    pb, eb = fixture_meshes[name_b]
    dim = pa.shape[1]
    total = 0.0
    for a, b in _pairs_closed(pa, ea, pb, eb):
        xa, xb = pa[ea[a]], pb[eb[b]]
        la, lb = ea.shape[1], eb.shape[1]
        if dim == 2:
            if la == 3 and lb == 4:
                out = O.intersect_tri_quad(xa, xb)
            elif la == 4 and lb == 3:
                out = O.intersect_tri_quad(xb, xa)
            elif la == 4 and lb == 4:
                out = O.intersect_quads(xa, xb)
            else:
                out = O.intersect_tris(xa, xb)
            total += sum(O.triangle_area(t) for t in out)
        else:
            if la == 4 and lb == 8:
                out = O.intersect_tet_hex(xa, xb)
            elif la == 8 and lb == 4:
                out = O.intersect_tet_hex(xb, xa)
            elif la == 8 and lb == 8:
                out = O.intersect_hexes(xa, xb)
            else:
                out = O.intersect_tets(xa, xb)
            total += sum(O.tet_volume_real(t) for t in out)
    return total


@pytest.mark.parametrize("names,expected", [
    (("square.1", "square.4"), 1.0),   # tri x quad
    (("square.3", "square.2"), 1.0),   # quad x tri
    (("square.3", "square.4"), 1.0),   # quad x quad
    (("cube.1", "cube.3"), 0.5),       # tet x hex
    (("cube.3", "cube.2"), 0.5),       # hex x tet
    (("cube.3", "cube.4"), 0.5),       # hex x hex
])
def test_reference_element_intersector_mixed(fixture_meshes, names, expected):
    # src/tests/test_element_intersector.F90: the known totals per mesh pair
    assert fequals(_sum_mixed(fixture_meshes, *names), expected)


def test_hex_planes_unit_cube_outward():
    n, c = O.hex_planes(CUBE)
    # faces y=0, z=1, x=1, z=0, x=0, y=1 with outward unit normals
    expect_n = np.array([[0, -1, 0], [0, 0, 1], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 1, 0]], float)
    expect_c = np.array([0, 1, 1, 0, 0, 1], float)
    assert np.array_equal(n, expect_n)
    assert np.array_equal(c, expect_c)


def _box(lo, hi):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    return lo + CUBE * (hi - lo)


def _box_overlap(alo, ahi, blo, bhi):
    v = Fraction(1)
    for d in range(3):
        v *= max(Fraction(0), min(Fraction(ahi[d]), Fraction(bhi[d])) - max(Fraction(alo[d]), Fraction(blo[d])))
    return v


def test_hexes_axis_aligned_boxes_exact():
    rng = np.random.default_rng(11)
    hits = 0
    for _ in range(200):
        alo = rng.uniform(0, 0.6, 3)
        ahi = alo + rng.uniform(0.1, 0.5, 3)
        blo = rng.uniform(0, 0.6, 3)
        bhi = blo + rng.uniform(0.1, 0.5, 3)
        out = O.intersect_hexes(_box(alo, ahi), _box(blo, bhi))
        vol = sum(O.tet_volume_real(t) for t in out)
        exact = float(_box_overlap(alo, ahi, blo, bhi))
        assert abs(vol - exact) <= 1e-15 + 1e-13 * exact
        hits += exact > 0
    assert hits > 50


def test_tet_hex_equals_sum_over_kuhn_tets():
    # tet clipped by a box's 6 planes == sum of tet-tet clips against the
    # box's 6 Kuhn tets (a different plane set and path)
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(300):
        lo = rng.uniform(0, 0.5, 3)
        hi = lo + rng.uniform(0.2, 0.5, 3)
        box = _box(lo, hi)
        tet = rng.uniform(0, 1, (4, 3))
        v6 = sum(O.tet_volume_real(t) for t in O.intersect_tet_hex(tet, box))
        vk = sum(O.tet_volume_real(t) for k in KUHN for t in O.intersect_tets(tet, box[list(k)]))
        scale = abs(O.tet_volume_real(tet))
        assert abs(v6 - vk) <= 1e-13 * max(scale, 1e-300)
        hits += v6 > 0
    assert hits > 50


def test_tet_hex_whole_tet_inside():
    box = _box([0, 0, 0], [1, 1, 1])
    tet = np.array([[0.1, 0.1, 0.1], [0.8, 0.2, 0.1], [0.3, 0.7, 0.2], [0.2, 0.3, 0.9]])
    out = O.intersect_tet_hex(tet, box)
    assert len(out) == 1
    assert np.array_equal(out[0], tet)


def _exact_poly_overlap_area(pa, pb):
    """|A ^ B| for a polygon A clipped by convex polygon B, in rationals."""
    pa = [(Fraction(x), Fraction(y)) for x, y in pa]
    pb = [(Fraction(x), Fraction(y)) for x, y in pb]
    n = len(pb)
    orient = sum(pb[i][0] * pb[(i + 1) % n][1] - pb[(i + 1) % n][0] * pb[i][1] for i in range(n))
    poly = pa
    for k in range(n):
        p, q = pb[k], pb[(k + 1) % n]

        def side(v):
            return ((q[0] - p[0]) * (v[1] - p[1]) - (q[1] - p[1]) * (v[0] - p[0])) * (1 if orient > 0 else -1)

        out = []
        for i in range(len(poly)):
            u, v = poly[i], poly[(i + 1) % len(poly)]
            su, sv = side(u), side(v)
            if su >= 0:
                out.append(u)
            if (su > 0 > sv) or (su < 0 < sv):
                t = su / (su - sv)
                out.append((u[0] + t * (v[0] - u[0]), u[1] + t * (v[1] - u[1])))
        poly = out
        if len(poly) < 3:
            return Fraction(0)
    a = sum(poly[i][0] * poly[(i + 1) % len(poly)][1] - poly[(i + 1) % len(poly)][0] * poly[i][1]
            for i in range(len(poly)))
    return abs(a) / 2


def _convex_quad(rng):
    """A random strictly convex quad, either orientation."""
    while True:
        c = rng.uniform(0.2, 0.8, 2)
        ang = np.sort(rng.uniform(0, 2 * np.pi, 4))
        r = rng.uniform(0.15, 0.5, 4)
        q = c + np.stack([r * np.cos(ang), r * np.sin(ang)], axis=1)
        e = np.roll(q, -1, axis=0) - q
        cross = e[:, 0] * np.roll(e, -1, axis=0)[:, 1] - e[:, 1] * np.roll(e, -1, axis=0)[:, 0]
        if np.all(cross > 1e-3):
            return q if rng.uniform() < 0.5 else q[::-1].copy()


def test_tri_quad_and_quads_vs_exact_rationals():
    rng = np.random.default_rng(21)
    hits = 0
    for _ in range(300):
        tri = rng.uniform(0, 1, (3, 2))
        qa, qb = _convex_quad(rng), _convex_quad(rng)
        a1 = sum(O.triangle_area(t) for t in O.intersect_tri_quad(tri, qb))
        assert abs(a1 - float(_exact_poly_overlap_area(tri, qb))) <= 1e-15
        a2 = sum(O.triangle_area(t) for t in O.intersect_quads(qa, qb))
        e2 = float(_exact_poly_overlap_area(qa, qb))
        assert abs(a2 - e2) <= 1e-15
        hits += e2 > 0
    assert hits > 100


# --------------------------------------------------------------------------
# 1D
# --------------------------------------------------------------------------

def test_intervals_closed_forms():
    assert O.interval_size([0.75, 0.25]) == 0.5
    assert O.intersect_intervals([0.0, 1.0], [0.5, 2.0]) == (1, pytest.approx([0.5, 1.0]))
    n, c = O.intersect_intervals([1.0, 0.0], [2.0, 0.5])   # reversed end points
    assert n == 1 and list(c) == [0.5, 1.0]
    assert O.intersect_intervals([0.0, 1.0], [1.0, 2.0])[0] == 0   # touching
    assert O.intersect_intervals([0.0, 1.0], [-2.0, 0.0])[0] == 0
    # overlap below 10 spacing(size) is dropped
    assert O.intersect_intervals([0.0, 1.0], [1.0 - 4e-16, 2.0])[0] == 0
    assert O.intersect_intervals([0.0, 1.0], [1.0 - 1e-14, 2.0])[0] == 1


def _strict_pairs_1d(pa, ea, pb, eb):
    a = np.sort(pa[ea, 0], axis=1)
    b = np.sort(pb[eb, 0], axis=1)
    hit = (b[None, :, 1] > a[:, None, 0]) & (b[None, :, 0] < a[:, None, 1])
    return set(zip(*np.nonzero(hit)))


def test_line_meshes_sum_and_completeness(fixture_meshes):
    # src/tests/test_interval_intersector.F90 (sum 1) and
    # test_intersection_finder_completeness_1d.F90 (per-element completeness)
    pa, ea = fixture_meshes["line.1"]
    pb, eb = fixture_meshes["line.2"]
    for (xa, xe), (ya, ye) in (((pa, ea), (pb, eb)), ((pb, eb), (pa, ea))):
        indptr, indices = O.sort_finder(xa, xe, ya, ye)
        total = 0.0
        for a in range(xe.shape[0]):
            size_a = O.interval_size(xa[xe[a], 0])
            acc = 0.0
            for b in indices[indptr[a]:indptr[a + 1]]:
                n, c = O.intersect_intervals(xa[xe[a], 0], ya[ye[b], 0])
                if n:
                    acc += O.interval_size(c)
            assert fequals(acc, size_a)
            total += acc
        assert fequals(total, 1.0)
        got = set(zip(*O.csr_to_pairs(indptr, indices)))
        assert got == _strict_pairs_1d(xa, xe, ya, ye)


def _line_mesh(rng, n, shuffle_nodes):
    x = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, n - 1)), [1.0]])
    enl = np.stack([np.arange(n), np.arange(1, n + 1)], axis=1).astype(np.int32)
    if shuffle_nodes:
        enl = enl[:, ::-1].copy()   # vertex order inside elements reversed
    return x[:, None], enl


@pytest.mark.parametrize("shuffle", [False, True])
@pytest.mark.parametrize("na,nb", [(1, 1), (2, 7), (37, 50), (300, 211)])
def test_sort_finder_equals_strict_brute_force_on_meshes(na, nb, shuffle):
    # on meshes (non-overlapping elements) the sweep finds exactly the strict
    # overlap set, whatever the element numbering
    rng = np.random.default_rng(na * 1000 + nb)
    pa, ea = _line_mesh(rng, na, False)
    pb, eb = _line_mesh(rng, nb, True)
    if shuffle:
        ea = np.ascontiguousarray(ea[rng.permutation(na)])
        eb = np.ascontiguousarray(eb[rng.permutation(nb)])
    indptr, indices = O.sort_finder(pa, ea, pb, eb)
    got = list(zip(*O.csr_to_pairs(indptr, indices)))
    assert len(got) == len(set(got))
    assert set(got) == _strict_pairs_1d(pa, ea, pb, eb)
    # each row lists B in ascending left end point (the sweep order)
    lb = np.minimum(pb[eb[:, 0], 0], pb[eb[:, 1], 0])
    for a in range(ea.shape[0]):
        row = indices[indptr[a]:indptr[a + 1]]
        assert np.all(np.diff(lb[row]) > 0)


@pytest.mark.parametrize("dim,n", [(2, 6), (3, 3)])
def test_cell_grid_tiles_the_unit_domain(dim, n):
    # the quad / hex generator of the mixed-type workloads: convex cells in
    # the reference's node order whose clips against a simplex mesh of the
    # same domain sum to exactly its measure
    from libsupermesh_amd import meshgen
    pq, eq = meshgen.cell_grid(dim, n, 3)
    ps, es = meshgen.split_grid(dim, n + 1, 4)
    if dim == 2:
        q = pq[eq]
        d = np.roll(q, -1, axis=1) - q
        cross = d[:, :, 0] * np.roll(d, -1, axis=1)[:, :, 1] - d[:, :, 1] * np.roll(d, -1, axis=1)[:, :, 0]
        assert (cross > 0).all()  # counter-clockwise and strictly convex
    else:
        h = pq[eq]
        assert np.array_equal(h[:, 1, 1:], h[:, 0, 1:]) and np.array_equal(h[:, 4, :2], h[:, 0, :2])
    meas = O.triangle_area if dim == 2 else O.tet_volume_real
    fn = O.intersect_tri_quad if dim == 2 else O.intersect_tet_hex
    total = 0.0
    for a, b in _pairs_closed(ps, es, pq, eq):
        total += sum(meas(t) for t in fn(ps[es[a]], pq[eq[b]]))
    assert abs(total - 1.0) < 1e-12
