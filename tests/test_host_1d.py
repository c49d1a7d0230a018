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


def _intervals(iv):
    """A 1D mesh of disjoint-node intervals: element e spans iv[e] (nodes 2e, 2e+1)."""
    x = np.array([v for pair in iv for v in pair], dtype=np.float64)[:, None]
    enl = np.arange(2 * len(iv), dtype=np.int32).reshape(-1, 2)
    return x, enl


# Worked by hand through the reference's sweep (Intersection_Finder.F90:811-866:
# A and B sorted by left end, one B cursor j; B skipped while max(b) <= min(a),
# the scan for a stops at min(b) >= max(a), and j steps back one after an
# element that found any).  Where intervals overlap within one mesh the sweep
# misses pairs that brute force finds; these cases fix which.  Every mesh has
# its elements in ascending left-end order within each base block of the
# reference's merge sort, where that sort is defined.
SWEEP_CASES = [
    # nested A: a1 = [1, 9] meets b0 = [2, 3] too, but j stays past it
    ("nested_a", [(0, 10), (1, 9)], [(2, 3), (4, 5)], [[0, 1], [1]]),
    # a long B met again only by the first A: j moved past it for good
    ("long_b", [(0, 1), (1, 2), (2, 3)], [(0, 3), (0.5, 0.7)], [[0, 1], [], []]),
    # j steps back one only: b0 = [1, 2.5] also meets a1 = [2, 5], unseen
    ("step_back_one", [(0, 3), (2, 5)], [(1, 2.5), (2.2, 2.4)], [[0, 1], [1]]),
    # the step back: b0 serves both A elements
    ("shared_b", [(0, 2), (2, 4)], [(1, 3)], [[0], [0]]),
    # touching ends are not overlaps (<= and >= in the sweep)
    ("touching", [(0, 1)], [(-1, 0), (1, 2)], [[]]),
    # eight B elements, sorted halves interleaved by the merge: the row is in
    # left-end order, ties (1, 2, 3) taking the first half's element first
    ("merge_order", [(-1, 10)], [(0, 0.5), (2, 2.5), (4, 4.5), (6, 6.5), (1, 1.5), (3, 3.5), (5, 5.5), (7, 7.5)],
     [[0, 4, 1, 5, 2, 6, 3, 7]]),
    ("merge_ties", [(-1, 10)], [(0, 0.5), (1, 1.5), (2, 2.5), (3, 3.5), (1, 1.25), (2, 2.25), (3, 3.25), (4, 4.5)],
     [[0, 1, 4, 2, 5, 3, 6, 7]]),
]


@pytest.mark.parametrize("name,ia,ib,rows", SWEEP_CASES, ids=[c[0] for c in SWEEP_CASES])
def test_sort_finder_sweep_worked_cases(name, ia, ib, rows):
    pa, ea = _intervals(ia)
    pb, eb = _intervals(ib)
    want_ptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    want_idx = np.array([b for r in rows for b in r], dtype=np.int64)
    for indptr, indices in (gpu.sort_intersection_finder(pa, ea, pb, eb), O.sort_finder(pa, ea, pb, eb)):
        assert np.array_equal(indptr, want_ptr), (name, indptr)
        assert np.array_equal(np.asarray(indices, dtype=np.int64), want_idx), (name, indices)
