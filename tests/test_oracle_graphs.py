"""Pins of the oracle's mesh adjacency (Graphs.F90, C_bindings.F90:176-430)
and advancing front finder (Intersection_Finder.F90:270-454), each against an
independent construction: numpy / scipy, closed forms of structured grids and
hand-worked discovery orders."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
from libsupermesh_amd import meshgen


def _meshes():
    out = []
    for dim in (2, 3):
        p, e = meshgen.split_grid(dim, 4 if dim == 2 else 2, 1)
        out.append((p.shape[0], e))
        p, e = meshgen.cell_grid(dim, 3, 1)
        out.append((p.shape[0], e))
    rng = np.random.default_rng(4)
    out.append((9, rng.integers(0, 9, (12, 3))))  # repeated nodes, unused nodes
    return out


@pytest.mark.parametrize("k", range(5))
def test_nelist_is_stable_sort_by_node(k):
    nn, enl = _meshes()[k]
    ptr, idx = O.mesh_nelist(nn, enl)
    flat = enl.reshape(-1)
    order = np.argsort(flat, kind="stable")
    assert np.array_equal(idx, order // enl.shape[1])
    assert np.array_equal(ptr, np.concatenate([[0], np.cumsum(np.bincount(flat, minlength=nn))]))


@pytest.mark.parametrize("k", range(5))
def test_nnlist_is_sorted_node_adjacency(k):
    nn, enl = _meshes()[k]
    ptr, idx = O.mesh_nnlist(nn, enl)
    loc = enl.shape[1]
    r, c = [], []
    for i in range(loc):
        for j in range(i + 1, loc):
            r += [enl[:, i], enl[:, j]]
            c += [enl[:, j], enl[:, i]]
    r, c = np.concatenate(r), np.concatenate(c)
    A = sp.coo_matrix((np.ones(len(r)), (r, c)), shape=(nn, nn)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    assert np.array_equal(ptr, A.indptr) and np.array_equal(idx, A.indices)


def _face_neighbours(enl, sloc):
    """Elements sharing at least sloc nodes, as sets (structured meshes)."""
    sets = [set(e) for e in enl.tolist()]
    by_node = {}
    for i, e in enumerate(enl.tolist()):
        for x in set(e):
            by_node.setdefault(x, []).append(i)
    out = []
    for i, s in enumerate(sets):
        cand = {j for x in s for j in by_node[x] if j != i}
        out.append({j for j in cand if len(s & sets[j]) >= sloc})
    return out


@pytest.mark.parametrize("k", range(4))
def test_eelist_sets_are_face_neighbours(k):
    nn, enl = _meshes()[k]
    loc = enl.shape[1]
    dim = {3: 2, 4: 2 if k == 1 else 3, 8: 3}[loc]
    simplex = loc == dim + 1
    sloc, nneigh = (loc - 1, loc) if simplex else (loc // 2, 2 * dim)
    ptr, idx = O.mesh_eelist(nn, enl, sloc, nneigh)
    ref = _face_neighbours(enl, sloc)
    for i in range(len(enl)):
        row = list(idx[ptr[i]:ptr[i + 1]])
        assert len(row) == len(set(row)) and set(row) == ref[i], i
        for j in row:  # symmetric
            assert i in idx[ptr[j]:ptr[j + 1]]


def test_eelist_discovery_order_hand_worked():
    # four triangles around node 4 plus e4 = (1, 0, 5) below edge 0-1.
    # e0 = (0, 1, 4): node 0 lists e3, e4 (one shared node each); node 1
    # lists e1 and e4 (e4 now shares 2 -> first); node 4 lists e1 (2 -> next),
    # e2, e3 (2 -> last): [4, 1, 3], not ascending.
    enl = np.array([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4], [1, 0, 5]])
    ptr, idx = O.mesh_eelist(6, enl, 2, 3)
    rows = [list(idx[ptr[i]:ptr[i + 1]]) for i in range(5)]
    assert rows == [[4, 1, 3], [0, 2], [1, 3], [0, 2], [0]]
    # the release build stops at nneigh
    ptr, idx = O.mesh_eelist(6, enl, 2, 1)
    assert [list(idx[ptr[i]:ptr[i + 1]]) for i in range(5)] == [[4], [0], [1], [0], [0]]


def test_transpose_is_scipy_transpose():
    rng = np.random.default_rng(7)
    A = sp.random(40, 25, density=0.2, format="csr", random_state=rng)
    A.sort_indices()
    ptr_t, idx_t = O.transpose_sparsity(25, A.indptr, A.indices)
    T = A.T.tocsr()
    T.sort_indices()
    assert np.array_equal(ptr_t, T.indptr) and np.array_equal(idx_t, T.indices)
    ptr2, idx2 = O.transpose_sparsity(40, ptr_t, idx_t)
    assert np.array_equal(ptr2, A.indptr) and np.array_equal(idx2, A.indices)


def _closed_sets(pa, ea, pb, eb):
    ip, ix = O.brute_force(pa, ea, pb, eb, closed=True)
    return [set(ix[ip[a]:ip[a + 1]].tolist()) for a in range(len(ea))]


@pytest.mark.parametrize("case", ["grid2", "grid3", "quads", "square", "cube"])
def test_advancing_front_sets_and_front_order(case, fixture_meshes):
    if case == "grid2":
        pa, ea, pb, eb = meshgen.config_meshes("tri_64_vs_50")
    elif case == "grid3":
        pa, ea = meshgen.split_grid(3, 4, 2)
        pb, eb = meshgen.split_grid(3, 5, 3, diagonal=1)
    elif case == "quads":
        pa, ea = meshgen.split_grid(2, 9, 2)
        pb, eb = meshgen.cell_grid(2, 7, 3)
    elif case == "square":
        pa, ea = fixture_meshes["square.1"]
        pb, eb = fixture_meshes["square.2"]
    else:
        pa, ea = fixture_meshes["cube.1"]
        pb, eb = fixture_meshes["cube.2"]
    ptr, idx = O.advancing_front(pa, ea, pb, eb)
    ref = _closed_sets(pa, ea, pb, eb)
    rows = [list(idx[ptr[a]:ptr[a + 1]]) for a in range(len(ea))]
    # every row: no repeats, inside the closed brute-force set, and holding
    # every B element that really intersects (the front may skip elements
    # whose boxes only touch A's box across a gap)
    clip = {(3, 3): O.intersect_tris, (4, 4): O.intersect_tets, (3, 4): O.intersect_tri_quad}
    fn = clip[(ea.shape[1], eb.shape[1])]
    for a, row in enumerate(rows):
        assert len(row) == len(set(row)) and set(row) <= ref[a], a
        for b in ref[a] - set(row):
            assert len(fn(pa[ea[a]], pb[eb[b]])) == 0, (a, b)
    # front order: element 0 is the seed (ascending brute force); every other
    # row starts with a face neighbour's row filtered to its own set, in that
    # neighbour's order, and every later entry is a B face neighbour of an
    # earlier one
    assert rows[0] == sorted(ref[0])
    assert len(idx) >= 0.99 * sum(len(r) for r in ref)
    dim, la, lb = pa.shape[1], ea.shape[1], eb.shape[1]
    sl = lambda loc: loc - 1 if loc == dim + 1 else loc // 2  # noqa: E731
    nb = lambda loc: loc if loc == dim + 1 else 2 * dim  # noqa: E731
    eptr_a, eidx_a = O.mesh_eelist(pa.shape[0], ea, sl(la), nb(la))
    eptr_b, eidx_b = O.mesh_eelist(pb.shape[0], eb, sl(lb), nb(lb))
    nbr_b = [set(eidx_b[eptr_b[b]:eptr_b[b + 1]].tolist()) for b in range(len(eb))]

    def grown_from(row, c):
        pre = [b for b in rows[c] if b in set(row)]
        if row[:len(pre)] != pre:
            return False
        return all(set(row[:i]) & nbr_b[row[i]] for i in range(len(pre), len(row)))

    for a in range(1, len(ea)):
        assert any(grown_from(rows[a], c) for c in eidx_a[eptr_a[a]:eptr_a[a + 1]]), a
