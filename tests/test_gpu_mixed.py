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
