"""Per-pair output counts of the oracle's clippers pinned by what the
mathematics fixes, independently of the clippers' own formulas.

* Triangles (Triangle_Intersection.F90:144-185,282-335): on generic random
  pairs the Sutherland-Hodgman polygon is the convex polygon A ^ B, whose
  vertices are computed here exactly (rational arithmetic) as the vertices of
  A inside B, of B inside A and the edge crossings; the reference fans it
  from its first vertex and drops no triangle whose area is far above the
  tolerance, so the count is (vertex count - 2).  Output vertices must sit on
  the exact polygon's vertices and the areas must sum to its area.
* Tetrahedra (Tetrahedron_Intersection.F90:137-179,244-436): on generic
  random pairs, count > 0 exactly when the exact overlap volume is positive
  (the overlap polyhedron is computed here as the vertices of the eight
  half-spaces' arrangement); output volumes sum to it and every output vertex
  lies in both tets.
* The branches decided at the last bit, each on dyadic inputs whose exact
  values fall on the branch, so that the float computation is exact and the
  count follows from the reference's comparison alone:
  - a vertex exactly on B's line, d1 == 0 (Triangle_Intersection.F90:321);
  - a tet vertex within epsilon = 2^-52 of an axis plane of B, abs(d) <
    epsilon (Tetrahedron_Intersection.F90:264-276), and at exactly epsilon;
  - a fan triangle of area exactly 10 * spacing, and 9 and 11 times
    spacing (the keep test area < tol, Triangle_Intersection.F90:174-183);
  - a tet of volume exactly 10 * spacing, and 9 and 11 times spacing (the
    keep test vol >= tol, Tetrahedron_Intersection.F90:150-176).
"""
import itertools
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle_lib as O


def spacing(x):
    """Fortran SPACING of a positive normal double, from its exponent alone:
    2^(e - 53) with x = m 2^e, 0.5 <= m < 1."""
    m, e = math.frexp(x)
    return math.ldexp(1.0, e - 53)


# --------------------------------------------------------------------------
# Exact convex geometry (rationals)
# --------------------------------------------------------------------------

def _cross2(o, a, b):
    return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])


def _inside_tri(p, t):
    """p in the closed triangle t (any orientation), exactly."""
    s = [_cross2(t[i], t[(i + 1) % 3], p) for i in range(3)]
    return all(x >= 0 for x in s) or all(x <= 0 for x in s)


def _seg_cross(p, q, r, s):
    """Intersection point of the segments pq and rs when they cross at one point."""
    d = (q[0] - p[0]) * (s[1] - r[1]) - (q[1] - p[1]) * (s[0] - r[0])
    if d == 0:
        return None
    t = ((r[0] - p[0]) * (s[1] - r[1]) - (r[1] - p[1]) * (s[0] - r[0])) / d
    u = ((r[0] - p[0]) * (q[1] - p[1]) - (r[1] - p[1]) * (q[0] - p[0])) / d
    if 0 <= t <= 1 and 0 <= u <= 1:
        return (p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1]))
    return None


def exact_tri_polygon(ta, tb):
    """Vertices of the convex polygon A ^ B, exact, in counter-clockwise order."""
    A = [(Fr(x), Fr(y)) for x, y in ta]
    B = [(Fr(x), Fr(y)) for x, y in tb]
    pts = {p for p in A if _inside_tri(p, B)} | {p for p in B if _inside_tri(p, A)}
    for i in range(3):
        for j in range(3):
            x = _seg_cross(A[i], A[(i + 1) % 3], B[j], B[(j + 1) % 3])
            if x is not None:
                pts.add(x)
    pts = list(pts)
    if len(pts) < 3:
        return pts
    cx = sum(p[0] for p in pts) / len(pts)
    cy = sum(p[1] for p in pts) / len(pts)
    pts.sort(key=lambda p: math.atan2(float(p[1] - cy), float(p[0] - cx)))
    # drop exactly collinear points (none on generic inputs)
    out = []
    for k in range(len(pts)):
        if _cross2(pts[k - 1], pts[k], pts[(k + 1) % len(pts)]) != 0:
            out.append(pts[k])
    return out


def _poly_area(P):
    return abs(sum(P[i][0] * P[(i + 1) % len(P)][1] - P[(i + 1) % len(P)][0] * P[i][1]
                   for i in range(len(P)))) / 2


def _tet_halfspaces(t):
    """Four (n, c) with n.x <= c inside, exact, for tet t (4 x 3 rationals)."""
    out = []
    for i in range(4):
        f = [t[j] for j in range(4) if j != i]
        u = [f[1][k] - f[0][k] for k in range(3)]
        v = [f[2][k] - f[0][k] for k in range(3)]
        n = [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
        c = sum(n[k] * f[0][k] for k in range(3))
        if sum(n[k] * t[i][k] for k in range(3)) > c:   # the opposite vertex is inside
            n = [-x for x in n]
            c = -c
        out.append((n, c))
    return out


def _solve3(rows, rhs):
    """Cramer's rule; None when singular."""
    (a, b, c), (d, e, f), (g, h, i) = rows
    det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)
    if det == 0:
        return None
    r0, r1, r2 = rhs
    x = (r0 * (e * i - f * h) - b * (r1 * i - f * r2) + c * (r1 * h - e * r2)) / det
    y = (a * (r1 * i - f * r2) - r0 * (d * i - f * g) + c * (d * r2 - r1 * g)) / det
    z = (a * (e * r2 - r1 * h) - b * (d * r2 - r1 * g) + r0 * (d * h - e * g)) / det
    return (x, y, z)


def exact_tet_overlap(ta, tb):
    """(volume, half-spaces) of the convex polyhedron A ^ B, exact: its vertices
    are the feasible points where three of the eight face planes meet; each
    face's vertices are ordered around the face and fanned to the centroid."""
    A = [[Fr(x) for x in v] for v in ta]
    B = [[Fr(x) for x in v] for v in tb]
    H = _tet_halfspaces(A) + _tet_halfspaces(B)
    verts = set()
    for i, j, k in itertools.combinations(range(8), 3):
        p = _solve3([H[i][0], H[j][0], H[k][0]], [H[i][1], H[j][1], H[k][1]])
        if p is not None and all(sum(n[q] * p[q] for q in range(3)) <= c for n, c in H):
            verts.add(p)
    verts = list(verts)
    if len(verts) < 4:
        return Fr(0), H
    g = [sum(v[q] for v in verts) / len(verts) for q in range(3)]
    vol = Fr(0)
    for n, c in H:
        face = [v for v in verts if sum(n[q] * v[q] for q in range(3)) == c]
        if len(face) < 3:
            continue
        fc = [sum(v[q] for v in face) / len(face) for q in range(3)]
        # order around the face centroid in the plane's own 2D frame
        nf = [float(x) for x in n]
        ax = np.array([1.0, 0, 0]) if abs(nf[0]) < 0.9 * np.linalg.norm(nf) else np.array([0, 1.0, 0])
        e1 = np.cross(nf, ax)
        e2 = np.cross(nf, e1)
        face.sort(key=lambda v: math.atan2(float(sum((v[q] - fc[q]) * e2[q] for q in range(3))),
                                           float(sum((v[q] - fc[q]) * e1[q] for q in range(3)))))
        for m in range(len(face)):
            a, b = face[m], face[(m + 1) % len(face)]
            u = [a[q] - g[q] for q in range(3)]
            v = [b[q] - g[q] for q in range(3)]
            w = [fc[q] - g[q] for q in range(3)]
            det = (u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0])
                   + u[2] * (v[0] * w[1] - v[1] * w[0]))
            vol += abs(det) / 6
    return vol, H


# --------------------------------------------------------------------------
# Generic pairs
# --------------------------------------------------------------------------

def _random_tri(rng):
    while True:
        t = rng.uniform(0, 1, (3, 2))
        if abs(_cross2(*[tuple(x) for x in t])) > 0.02:
            return t


def test_tri_counts_generic_pairs_vs_exact_polygon():
    rng = np.random.default_rng(20)
    checked = nonempty = 0
    seen_counts = set()
    while checked < 10000:
        ta, tb = _random_tri(rng), _random_tri(rng)
        P = exact_tri_polygon(ta, tb)
        # genericity: no vertex of one within 1e-7 of the other's lines, and
        # every fan triangle from any apex far above the tolerance
        A = [(Fr(x), Fr(y)) for x, y in ta]
        B = [(Fr(x), Fr(y)) for x, y in tb]
        near = False
        for T, U in ((A, B), (B, A)):
            for i in range(3):
                p, q = U[i], U[(i + 1) % 3]
                ln = math.hypot(float(q[0] - p[0]), float(q[1] - p[1]))
                near |= any(abs(float(_cross2(p, q, v))) / ln < 1e-7 for v in T)
        if len(P) >= 3:
            near |= any(abs(float(_cross2(P[a], P[b], P[c]))) < 1e-9
                        for a, b, c in itertools.combinations(range(len(P)), 3))
        if near:
            continue
        checked += 1
        out = O.intersect_tris(ta, tb)
        expect = len(P) - 2 if len(P) >= 3 else 0
        assert len(out) == expect, (ta, tb)
        seen_counts.add(expect)
        if expect:
            nonempty += 1
            assert abs(sum(O.triangle_area(t) for t in out) - float(_poly_area(P))) <= 1e-14
            exact_pts = np.array([[float(x), float(y)] for x, y in P])
            d = np.linalg.norm(out.reshape(-1, 1, 2) - exact_pts[None], axis=2).min(axis=1)
            assert d.max() <= 1e-12
    assert nonempty > 2000 and {0, 1, 2, 3, 4} <= seen_counts


def test_tet_nonempty_iff_exact_overlap_volume():
    rng = np.random.default_rng(21)
    checked = nonempty = 0
    while checked < 500:
        ta = rng.uniform(0, 1, (4, 3))
        tb = rng.uniform(0, 1, (4, 3))
        if abs(np.linalg.det(ta[1:] - ta[0])) < 0.02 or abs(np.linalg.det(tb[1:] - tb[0])) < 0.02:
            continue
        vol, H = exact_tet_overlap(ta, tb)
        if 0 < vol < 1e-9:
            continue   # near-tangent: not generic
        # genericity: no vertex of either tet within 1e-9 of a face plane of the other
        near = False
        for (n, c), T in zip(H, [tb] * 4 + [ta] * 4):
            nf = math.sqrt(sum(float(x) ** 2 for x in n))
            near |= any(abs(float(sum(n[q] * Fr(v[q]) for q in range(3)) - c)) / nf < 1e-9 for v in T)
        if near:
            continue
        checked += 1
        out = O.intersect_tets(ta, tb)
        assert (len(out) > 0) == (vol > 0), (ta, tb)
        got = sum(O.tet_volume_real(t) for t in out)
        assert abs(got - float(vol)) <= 1e-14
        for t in out:
            for v in t:
                for n, c in H:
                    nf = math.sqrt(sum(float(x) ** 2 for x in n))
                    assert float(sum(n[q] * Fr(v[q]) for q in range(3)) - c) / nf <= 1e-13
        nonempty += vol > 0
    assert nonempty > 100


# --------------------------------------------------------------------------
# Branches decided at the last bit
# --------------------------------------------------------------------------

def test_tri_vertex_exactly_on_line():
    """P = (1, 0) of A lies exactly on B's edge line y = 0 (d1 == 0); its
    neighbour Q is outside and R inside.  A ^ B is the triangle (P, X, R), X =
    (1.5, 0) on QR, of area 1/4: one output triangle.  Treating d1 == 0 as
    outside would drop P and leave two points, i.e. no output."""
    B = np.array([[0.0, 0.0], [4.0, 0.0], [0.0, 4.0]])
    A = np.array([[1.0, 0.0], [2.0, -1.0], [1.0, 1.0]])
    P = exact_tri_polygon(A, B)
    assert float(_poly_area(P)) == 0.25 and len(P) == 3
    out = O.intersect_tris(A, B)
    assert len(out) == 1
    assert np.array_equal(out[0], [[1.0, 0.0], [1.5, 0.0], [1.0, 1.0]])
    # touching at the single point P: nothing
    A2 = np.array([[1.0, 0.0], [2.0, -1.0], [0.5, -1.0]])
    assert len(O.intersect_tris(A2, B)) == 0
    # P on the line, both neighbours inside: A unchanged
    A3 = np.array([[1.0, 0.0], [2.0, 1.0], [1.0, 1.0]])
    out = O.intersect_tris(A3, B)
    assert len(out) == 1 and np.array_equal(out[0], A3)


def _band_case(zv1):
    """B = the unit right tet (face 1 on z = 0, outward normal exactly
    (0, 0, -1), so the distance of a point to it is exactly -z).  A: V1 at
    height zv1, V2 inside, V3 and V4 outside through z = 0."""
    B = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    h = 2.0 ** -10
    # distances 2h, -2h, -6h from the plane: split weights 1/2 and 1/4, exact
    A = np.array([[0.25, 0.25, zv1], [0.25 + h, 0.25, 2 * h], [0.25, 0.25 + h, -2 * h],
                  [0.25 + h, 0.25 + h, -6 * h]])
    return A, B


@pytest.mark.parametrize("zv1", [2.0 ** -54, -(2.0 ** -54), 2.0 ** -52 * (1 - 2.0 ** -53),
                                 -(2.0 ** -52 * (1 - 2.0 ** -53))])
def test_tet_vertex_within_epsilon_of_plane(zv1):
    """|d| < 2^-52: V1 is classified on the plane, so with one vertex inside
    and two outside the part kept is the single tet (V1, V2, X3, X4), X on the
    plane, with V1 left exactly where it is."""
    A, B = _band_case(zv1)
    out = O.intersect_tets(A, B)
    assert len(out) == 1
    assert any(np.array_equal(v, A[0]) for v in out[0])
    assert any(np.array_equal(v, A[1]) for v in out[0])
    assert np.sum(out[0][:, 2] == 0.0) == 2


@pytest.mark.parametrize("zv1", [2.0 ** -52, 2.0 ** -51])
def test_tet_vertex_at_epsilon_from_plane(zv1):
    """|d| = 2^-52 is not < epsilon: V1 is inside, the kept part is a wedge
    (six vertices) that no single tet of its vertices covers, and the pieces
    sum to its exact volume."""
    A, B = _band_case(zv1)
    out = O.intersect_tets(A, B)
    assert len(out) >= 2
    vol, _ = exact_tet_overlap(A, B)
    # coordinates near 0.25 and edges of 2^-10: relative rounding ~1e-13
    assert abs(sum(O.tet_volume_real(t) for t in out) - float(vol)) <= 1e-12 * float(vol)


def test_tet_vertex_at_epsilon_outside_plane():
    """d = +2^-52 is not within the band either: V1 is outside, so with one
    vertex inside (V2) the kept part is the tet (V2, X1, X3, X4) and V1 is
    moved to the plane."""
    A, B = _band_case(-(2.0 ** -52))
    out = O.intersect_tets(A, B)
    assert len(out) == 1
    assert not any(np.array_equal(v, A[0]) for v in out[0])
    assert any(np.array_equal(v, A[1]) for v in out[0])


@pytest.mark.parametrize("k,kept", [(9, False), (10, True), (11, True)])
def test_tri_fan_triangle_at_the_tolerance(k, kept):
    """A = (0,0), (1,0), (0,1) (area 1/2, spacing 2^-53, tol = 10 * 2^-53 as
    B is larger); B's edge cuts A's edge V2V3 at fraction f = k 2^-52 from V2
    and V3V1 at its midpoint, all exactly.  The polygon is (V1, V2, X, Y);
    the fan triangle (V1, V2, X) has area f/2 = k 2^-53 exactly: kept unless
    below 10 * 2^-53."""
    f = k * 2.0 ** -52
    Z = (0.0, 0.5)
    P1 = (2.0 - 2.0 * f, 2.0 * f - 0.5)          # Z + 2 (X - Z), X = (1 - f, f)
    W = (-8.0, -8.0)
    A = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    # B's first edge Z -> P1 is the cut; distances to it from Z are exact:
    # -(1 - f) at V1, -f at V2, 1 - f at V3
    B = np.array([Z, P1, W])
    assert 10.0 * min(spacing(0.5), spacing(O.triangle_area(B))) == 10 * 2.0 ** -53
    X = np.array([1.0 - f, f])
    thin = np.array([[0.0, 0.0], [1.0, 0.0], X])
    big = np.array([[0.0, 0.0], X, [0.0, 0.5]])
    assert O.triangle_area(thin) == k * 2.0 ** -53
    P = exact_tri_polygon(A, B)
    assert len(P) == 4 and _poly_area(P) == Fr(f) / 2 + (1 - Fr(f)) / 4
    out = O.intersect_tris(A, B)
    if kept:
        assert len(out) == 2 and np.array_equal(out[0], thin) and np.array_equal(out[1], big)
    else:
        assert len(out) == 1 and np.array_equal(out[0], big)


@pytest.mark.parametrize("k,kept", [(9, False), (10, True), (11, True)])
def test_tet_at_the_tolerance(k, kept):
    """A's apex P lies just inside B's top face z = c (normal exactly (0,0,1))
    and its three other vertices far above it, inside B's other three planes:
    the kept part is the apex tet (P, X1, X2, X3), Xi at fraction s = 2^-17
    along each edge, exactly.  Its volume s^3 vol(A) is set to k * 2^-60 and
    B's volume to 2^-8 .. 2^-7 (spacing 2^-60), so the tolerance is exactly
    10 * 2^-60: kept unless below it (vol >= tol keeps)."""
    s = 2.0 ** -17
    H = 2.0                       # apex to top vertices, along z
    c = 0.5                       # B's top face
    # B: an inverted right tet, top face z = c (edges along x and y), bottom
    # vertex below; its volume L^3 / 6 with L = 5/16 lies in [2^-8, 2^-7).
    # Above the top face its other three planes bound a widening cone.
    L = 5.0 / 16
    x0, y0 = -0.09375, -0.09375
    B = np.array([[x0, y0, c], [x0 + L, y0, c], [x0, y0 + L, c], [x0, y0, c - L]])
    P = np.array([x0 + 0.0625, y0 + 0.0625, c - s * H])
    # top triangle at z = P_z + H with vol(A) = a b H / 6 = k 2^-9
    a = 2.0 ** -2
    b = k * 2.0 ** -9 * 6 / (a * H)     # = 3k 2^-7, exact
    top = P[2] + H
    A = np.array([P, [P[0], P[1], top], [P[0] + a, P[1], top], [P[0], P[1] + b, top]])
    assert O.tet_volume_tet(A) == k * 2.0 ** -9
    vb = O.tet_volume_tet(B)
    assert 2.0 ** -8 <= vb < 2.0 ** -7 and spacing(vb) == 2.0 ** -60
    assert spacing(O.tet_volume_tet(A)) > 2.0 ** -60
    # A's top vertices are outside B only through the top face
    n, cc = O.get_planes_tet(B)
    d = A @ n.T - cc
    top_face = int(np.argmax(np.abs(n[:, 2])))
    assert np.array_equal(n[top_face], [0.0, 0.0, 1.0]) or np.array_equal(n[top_face], [0.0, 0.0, -1.0])
    others = [i for i in range(4) if i != top_face]
    assert np.all(d[:, others] < -1e-6)
    out = O.intersect_tets(A, B)
    if kept:
        assert len(out) == 1
        assert O.tet_volume_tet(out[0]) == k * 2.0 ** -60
    else:
        assert len(out) == 0
