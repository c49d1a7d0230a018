"""The library's 1D entry points (host code: intervals and the sort finder)
against the oracle, bit for bit.  They touch no GPU, so they run on CPU."""
import numpy as np
import pytest

import oracle_lib as O
from libsupermesh_amd import gpu


def test_interval_functions_match_oracle():
    rng = np.random.default_rng(2)
    for _ in range(3000):
        a = rng.uniform(-1, 1, 2)
        b = rng.uniform(-1, 1, 2)
        if rng.uniform() < 0.2:
            b[0] = a[1] + rng.choice([0.0, 1e-16, -1e-16, 3e-15])  # near-touching
        assert gpu.interval_size(a) == O.interval_size(a)
        n, c = gpu.intersect_intervals(a, b)
        n_ref, c_ref = O.intersect_intervals(a, b)
        assert n == n_ref
        if n:
            assert np.array_equal(c, c_ref)


def _mesh(rng, n, order):
    x = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, n - 1)), [1.0]])
    enl = np.stack([np.arange(n), np.arange(1, n + 1)], axis=1)
    if order == "reversed_vertices":
        enl = enl[:, ::-1]
    elif order == "shuffled_elements":
        enl = enl[rng.permutation(n)]
    elif order == "shuffled_nodes":
        perm = rng.permutation(n + 1)
        inv = np.argsort(perm)
        x = x[perm]
        enl = inv[enl]
    return x[:, None], np.ascontiguousarray(enl)


@pytest.mark.parametrize("order", ["sorted", "reversed_vertices", "shuffled_elements", "shuffled_nodes"])
@pytest.mark.parametrize("na,nb", [(1, 1), (5, 3), (64, 77), (500, 333)])
def test_sort_finder_matches_oracle(na, nb, order):
    rng = np.random.default_rng(na + 7 * nb)
    pa, ea = _mesh(rng, na, order)
    pb, eb = _mesh(rng, nb, order)
    indptr, indices = gpu.sort_intersection_finder(pa, ea, pb, eb)
    r_indptr, r_indices = O.sort_finder(pa, ea, pb, eb)
    assert np.array_equal(indptr, r_indptr)
    assert np.array_equal(indices, r_indices)


def test_sort_finder_reference_line_meshes(fixture_meshes):
    pa, ea = fixture_meshes["line.1"]
    pb, eb = fixture_meshes["line.2"]
    indptr, indices = gpu.sort_intersection_finder(pa, ea, pb, eb)
    r_indptr, r_indices = O.sort_finder(pa, ea, pb, eb)
    assert np.array_equal(indptr, r_indptr) and np.array_equal(indices, r_indices)
    total = 0.0
    for a in range(ea.shape[0]):
        for b in indices[indptr[a]:indptr[a + 1]]:
            n, c = gpu.intersect_intervals(pa[ea[a], 0], pb[eb[b], 0])
            total += gpu.interval_size(c) if n else 0.0
    assert abs(total - 1.0) < 1e-14


def test_sort_finder_errors():
    x = np.array([[0.0], [1.0]])
    e = np.array([[0, 1]])
    with pytest.raises(gpu.GpuError, match="dimension"):
        gpu.sort_intersection_finder(np.zeros((2, 2)), e, np.zeros((2, 2)), e)
    with pytest.raises(gpu.GpuError, match="element type"):
        gpu.sort_intersection_finder(x, np.array([[0, 1, 1]]), x, np.array([[0, 1, 1]]))
    with pytest.raises(gpu.GpuError, match="out of range"):
        gpu.sort_intersection_finder(x, np.array([[0, 2]]), x, e)
    # a failed set_input leaves no result to query
    with pytest.raises(gpu.GpuError, match="before set_input"):
        n = gpu.ctypes.c_long(0)
        gpu.lib().libsupermesh_sort_intersection_finder_query_output(gpu.ctypes.byref(n))
        gpu._raise_if_error()
