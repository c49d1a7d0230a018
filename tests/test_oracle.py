"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each test is chosen so a plausible restatement mistake fails it: a wrong
NORM2 / dot-product rounding, a swapped split case or vertex order, a wrong
tolerance, or a finder predicate off by one comparison.
"""
import json
import os
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O
from libsupermesh_amd import meshgen

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(np.float64).eps


def fequals(a, b):
    """The reference tests' float comparison (src/Unittest_Tools.F90:179-195)."""
    tol = 100.0 * max(EPS, O.spacing(a), O.spacing(b))
    return abs(a - b) < tol


def stats():
    with open(os.path.join(HERE, "golden", "reference_stats.json")) as f:
        return json.load(f)


# --------------------------------------------------------------------------
# Fortran intrinsics, bit-compared with the reference's compiler
# --------------------------------------------------------------------------

@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not installed")
def test_intrinsics_bitwise_vs_amdflang(tmp_path):
    exe = tmp_path / "probe"
    subprocess.run(["amdflang", "-O2", "-o", str(exe),
                    os.path.join(HERE, "fortran", "probe_intrinsics.f90")], check=True)
    rng = np.random.default_rng(7)
    n = 20000
    v = rng.standard_normal((n, 6)) * np.exp(rng.uniform(-8, 8, (n, 1)))
    v[:100, 1] = 0.0          # zero components exercise NORM2's running maximum
    v[100:200, 0] = 0.0
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(v.tobytes())
    subprocess.run([str(exe), str(fin), str(fout)], check=True)
    r = np.fromfile(fout, dtype=np.float64).reshape(n, 7)
    for i in range(n):
        a, b = v[i, :3].copy(), v[i, 3:].copy()
        assert O.norm2(a) == r[i, 0]
        assert O.dot3(a, b) == r[i, 1]
        assert O.spacing(a[0]) == r[i, 2]
        c = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])
        assert np.array_equal(c / O.norm2(c), r[i, 3:6])


def test_norm2_is_not_the_naive_form():
    # The max-scaled NORM2 and sqrt(x*x+y*y+z*z) differ on a sizeable fraction
    # of inputs, so the bit test above really selects the form.
    rng = np.random.default_rng(1)
    v = rng.standard_normal((2000, 3))
    diff = sum(O.norm2(x) != np.sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) for x in v)
    assert diff > 100


def test_spacing_known_values():
    assert O.spacing(1.0) == 2.0 ** -52
    assert O.spacing(0.75) == 2.0 ** -53
    assert O.spacing(-3.0) == 2.0 ** -51
    assert O.spacing(0.0) == np.finfo(np.float64).tiny
    assert O.spacing(1e-310) == np.finfo(np.float64).tiny


# --------------------------------------------------------------------------
# Reference known answers on the reference's fixture meshes
# --------------------------------------------------------------------------

@pytest.mark.parametrize("key", ["triangle_0_01 x square_0_01", "cube_0_05 x pyramid_0_05"])
def test_reference_fixture_counts(fixture_meshes, key):
    ref = stats()[key]
    pa, ea = fixture_meshes[ref["a"]]
    pb, eb = fixture_meshes[ref["b"]]
    st, s, _ = O.supermesh(pa, ea, pb, eb, finder=2, nthreads=8)
    assert st["candidates"] == ref["candidates_closed"]
    assert st["nonempty"] == ref["nonempty"]
    assert st["elements"] == ref["elements"]
    assert st["max_per_pair"] == ref["max_per_pair"]
    assert s == ref["sum"]


def test_reference_tet_sum_other_orientation(fixture_meshes):
    pa, ea = fixture_meshes["pyramid_0_05"]
    pb, eb = fixture_meshes["cube_0_05"]
    st, s, _ = O.supermesh(pa, ea, pb, eb, finder=0, nthreads=8)
    assert st["candidates"] == 600057
    assert fequals(s, 1.0 / 3.0)


def test_reference_integrals(fixture_meshes):
    ref = stats()
    r2 = ref["integral_2d"]
    pa, ea = fixture_meshes[r2["a"]]
    pb, eb = fixture_meshes[r2["b"]]
    rows, cols = O.csr_to_pairs(*O.tree_finder(pa, ea, pb, eb))
    area, integral = O.inner_product(pa, ea, pa[:, 0], pb, eb, pb[:, 1], rows, cols)
    assert fequals(area, r2["area"])
    assert fequals(integral, r2["value"])
    r3 = ref["integral_3d"]
    pa, ea = fixture_meshes[r3["a"]]
    pb, eb = fixture_meshes[r3["b"]]
    rows, cols = O.csr_to_pairs(*O.tree_finder(pa, ea, pb, eb))
    vol, integral = O.inner_product(pa, ea, pa[:, 0] + pa[:, 2], pb, eb, pb[:, 1], rows, cols)
    assert fequals(vol, r3["volume"])
    assert fequals(integral, r3["value"])


def test_element_intersector_cubes(fixture_meshes):
    # src/tests/test_element_intersector.F90: cube.1 x cube.2 tet-tet, sum = 1
    pa, ea = fixture_meshes["cube.1"]
    pb, eb = fixture_meshes["cube.2"]
    _, s, _ = O.supermesh(pa, ea, pb, eb, finder=0)
    assert fequals(s, 1.0)
    pa, ea = fixture_meshes["square.1"]
    pb, eb = fixture_meshes["square.2"]
    _, s, _ = O.supermesh(pa, ea, pb, eb, finder=0)
    assert fequals(s, 1.0)


# --------------------------------------------------------------------------
# Finders
# --------------------------------------------------------------------------

def _numpy_pairs(pa, ea, pb, eb, closed):
    ba = pa[ea]
    bb = pb[eb]
    amin, amax = ba.min(axis=1), ba.max(axis=1)
    bmin, bmax = bb.min(axis=1), bb.max(axis=1)
    out = []
    for a0 in range(0, ea.shape[0], 512):
        lo, hi = amin[a0:a0 + 512, None, :], amax[a0:a0 + 512, None, :]
        if closed:
            hit = np.all((bmax[None] >= lo) & (bmin[None] <= hi), axis=2)
        else:
            hit = np.all((bmax[None] > lo) & (bmin[None] < hi), axis=2)
        r, c = np.nonzero(hit)
        out.append(np.stack([r + a0, c], axis=1))
    return np.concatenate(out)


def _pairs(indptr, indices):
    rows, cols = O.csr_to_pairs(indptr, indices)
    p = np.stack([rows, cols], axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


@pytest.mark.parametrize("closed", [False, True])
def test_brute_force_matches_numpy_predicate(closed):
    pa, ea, pb, eb = meshgen.config_meshes("tri_64_vs_50")
    got = _pairs(*O.brute_force(pa, ea, pb, eb, closed=closed))
    want = _numpy_pairs(pa, ea, pb, eb, closed)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("names", [("triangle_0_01", "square_0_01"), ("cube.1", "cube.2"),
                                   ("cube_0_05", "pyramid_0_05")])
def test_tree_equals_strict_brute_force(fixture_meshes, names):
    pa, ea = fixture_meshes[names[0]]
    pb, eb = fixture_meshes[names[1]]
    tree = _pairs(*O.tree_finder(pa, ea, pb, eb))
    brute = _pairs(*O.brute_force(pa, ea, pb, eb, closed=False))
    assert np.array_equal(tree, brute)


def test_tree_on_touching_grid():
    # Unjittered grids: many boxes touch exactly, so strict and closed differ.
    pa, ea = meshgen.split_grid(2, 8, 0, alpha=0.0, diagonal=0)
    pb, eb = meshgen.split_grid(2, 4, 0, alpha=0.0, diagonal=1)
    tree = _pairs(*O.tree_finder(pa, ea, pb, eb))
    strict = _pairs(*O.brute_force(pa, ea, pb, eb, closed=False))
    closed = _pairs(*O.brute_force(pa, ea, pb, eb, closed=True))
    assert np.array_equal(tree, strict)
    assert len(closed) > len(strict)


def test_tree_deep_split():
    # More than 256 elements force leaf splits (max_n = 256), recursively.
    pa, ea = meshgen.split_grid(3, 6, 5)
    pb, eb = meshgen.split_grid(3, 7, 6)
    tree = _pairs(*O.tree_finder(pa, ea, pb, eb))
    brute = _pairs(*O.brute_force(pa, ea, pb, eb, closed=False))
    assert np.array_equal(tree, brute)


@pytest.mark.parametrize("closed", [False, True])
@pytest.mark.parametrize("case", ["tri_64_vs_50", "fixture_2d", "fixture_3d", "touching", "far", "tiny",
                                  "mixed_sizes"])
def test_grid_finder_equals_brute_force(fixture_meshes, case, closed):
    """The cell-pruned finder used for full-size checks returns exactly the
    all-pairs set, rows ascending, on jittered and exact-touching grids,
    meshes far from the origin, tiny meshes, and A and B of very different
    element sizes."""
    if case == "tri_64_vs_50":
        pa, ea, pb, eb = meshgen.config_meshes("tri_64_vs_50")
    elif case == "fixture_2d":
        (pa, ea), (pb, eb) = fixture_meshes["triangle_0_01"], fixture_meshes["square_0_01"]
    elif case == "fixture_3d":
        (pa, ea), (pb, eb) = fixture_meshes["cube.1"], fixture_meshes["pyramid_0_05"]
    elif case == "touching":
        pa, ea = meshgen.split_grid(3, 6, 0, alpha=0.0, diagonal=0)
        pb, eb = meshgen.split_grid(3, 3, 0, alpha=0.0, diagonal=1)
    elif case == "far":
        pa, ea = meshgen.split_grid(2, 20, 1)
        pb, eb = meshgen.split_grid(2, 17, 2, mesh_id=1, diagonal=1)
        pa, pb = pa * 3.0 + 1e6, pb * 3.0 + 1e6
    elif case == "tiny":
        pa, ea = meshgen.split_grid(3, 4, 1)
        pb, eb = meshgen.split_grid(3, 5, 2, mesh_id=1)
        pa, pb = pa * 1e-9, pb * 1e-9
    else:
        pa, ea = meshgen.split_grid(2, 40, 1)
        pb, eb = meshgen.split_grid(2, 3, 2, mesh_id=1, diagonal=1)
    ip, ix = O.grid_finder(pa, ea, pb, eb, closed=closed, nthreads=3)
    bp, bx = O.brute_force(pa, ea, pb, eb, closed=closed)
    assert np.array_equal(ip, bp)
    assert np.array_equal(ix, bx)   # brute force rows are ascending too


# --------------------------------------------------------------------------
# Clippers against exact arithmetic and invariants
# --------------------------------------------------------------------------

def _exact_tri_overlap_area(ta, tb):
    ta = [(Fraction(x), Fraction(y)) for x, y in ta]
    tb = [(Fraction(x), Fraction(y)) for x, y in tb]
    orient = (tb[1][0] - tb[0][0]) * (tb[2][1] - tb[0][1]) - (tb[1][1] - tb[0][1]) * (tb[2][0] - tb[0][0])
    poly = ta
    for k in range(3):
        p, q = tb[k], tb[(k + 1) % 3]

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


def test_tri_clip_vs_exact_rationals():
    rng = np.random.default_rng(3)
    hits = 0
    for _ in range(400):
        ta = rng.uniform(0, 1, (3, 2))
        tb = rng.uniform(0, 1, (3, 2))
        out = O.intersect_tris(ta, tb)
        area = sum(O.triangle_area(t) for t in out)
        exact = float(_exact_tri_overlap_area(ta, tb))
        assert abs(area - exact) <= 1e-15  # coordinates are O(1)
        hits += exact > 0
    assert hits > 100


def test_tri_self_intersection_and_fan():
    t = np.array([[0.1, 0.2], [0.9, 0.3], [0.4, 0.8]])
    out = O.intersect_tris(t, t)
    assert len(out) == 1
    assert np.array_equal(out[0], t)
    # a triangle cut by a big one across two corners: hexagon -> 4 fan tris
    big = np.array([[-1.0, 0.5], [2.0, -0.9], [2.0, 2.0]])
    small = np.array([[0.0, 0.0], [1.0, 0.0], [0.5, 1.0]])
    out = O.intersect_tris(small, big)
    exact = float(_exact_tri_overlap_area(small, big))
    assert abs(sum(O.triangle_area(x) for x in out) - exact) < 1e-15


def test_tet_self_intersection(fixture_meshes):
    pos, enl = fixture_meshes["tet"]
    t = pos[enl[0]]
    out = O.intersect_tets(t, t)
    assert len(out) == 1
    assert np.array_equal(out[0], t)


def test_tet_completeness(fixture_meshes):
    # Sum over A of |a ^ b| equals |b| for every B element inside A's domain
    # (src/tests/test_intersection_finder_completeness_3d.F90:115-136).
    pa, ea = fixture_meshes["cube.1"]
    pb, eb = fixture_meshes["cube.2"]
    ip, ix = O.brute_force(pb, eb, pa, ea, closed=False)  # rows = B elements
    rb, ra = O.csr_to_pairs(ip, ix)
    counts, meas = O.clip_pairs(pa, ea, pb, eb, ra, rb)
    per_b = np.zeros(eb.shape[0])
    np.add.at(per_b, rb, meas)
    vol_b = np.array([O.tet_volume_real(pb[e]) for e in eb])
    assert np.allclose(per_b, vol_b, rtol=1e-12, atol=1e-15)


def test_tet_volume_formulas_differ_in_rounding():
    # _real (edges from v1, times 1/6) and _tet (edges from v4, divided by 6)
    # are both needed: they round differently.
    rng = np.random.default_rng(5)
    diff = 0
    for _ in range(2000):
        t = rng.uniform(0, 1, (4, 3))
        a, b = O.tet_volume_real(t), O.tet_volume_tet(t)
        assert abs(a - b) <= 1e-16  # coordinates are O(1)
        diff += a != b
    assert diff > 100


def test_tet_clip_translation_of_unit_tet():
    # Unit right tet against itself shifted by s along x: overlap is the
    # similar tet scaled by (1-s), volume (1-s)^3 / 6.
    t = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    for s in (0.25, 0.5, 0.125):
        u = t + np.array([s, 0, 0])
        out = O.intersect_tets(t, u)
        vol = sum(O.tet_volume_real(x) for x in out)
        assert abs(vol - (1 - s) ** 3 / 6) < 1e-15


# --------------------------------------------------------------------------
# Mixed mass
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dim", [2, 3])
def test_mass_exact_for_linear_fields(dim):
    if dim == 2:
        pa, ea, pb, eb = meshgen.config_meshes("tri_64_vs_50")
    else:
        pa, ea = meshgen.split_grid(3, 6, 11)
        pb, eb = meshgen.split_grid(3, 5, 12, diagonal=1)
    rows, cols = O.csr_to_pairs(*O.brute_force(pa, ea, pb, eb))
    blocks = O.mass_pairs(pa, ea, pb, eb, rows, cols)
    assert abs(blocks.sum() - 1.0) < 1e-13          # 1^T M 1 = |domain|
    f = pa[:, 0][ea[rows]]
    g = pb[:, 1][eb[cols]]
    fMg = np.einsum("pi,pij,pj->", f, blocks, g)
    assert abs(fMg - 0.25) < 1e-13                  # int_[0,1]^d x y = 1/4
