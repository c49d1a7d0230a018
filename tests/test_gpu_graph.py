"""The count pass recorded as one hipGraph (LIBSUPERMESH_GPU_GRAPH), in the
launch shape the context measured fastest: every replay gives the plain
count pass's results bit for bit -- per-pair counts and measures, the
compensated sum and a batch's per-mesh sums -- with the caller's own eager
work and synchronisations between the replays, and after the meshes move
(positions rewritten in place, so the recorded graph reads the new
coordinates).  The batched call replaces the caller loop of the reference
manual (doc/manual.tex:2064-2082) repeated over many mesh pairs."""
import numpy as np
import pytest

from libsupermesh_amd import gpu, meshgen

pytestmark = pytest.mark.gpu

SIZES = {2: [(3, 5), (12, 7), (1, 1), (9, 4), (20, 20), (2, 31), (16, 3)],
         3: [(2, 3), (4, 5), (1, 1), (6, 4), (1, 7), (5, 5)]}


def _batch(dim, seed):
    pairs = [(meshgen.split_grid(dim, na, seed, mesh_id=2 * k, diagonal=0),
              meshgen.split_grid(dim, nb, seed, mesh_id=2 * k + 1, diagonal=1))
             for k, (na, nb) in enumerate(SIZES[dim])]
    return gpu.DeviceMeshBatch([p[0] for p in pairs]), gpu.DeviceMeshBatch([p[1] for p in pairs])


def _outputs(r, br, dim):
    out = gpu.fetch(r, dim, ("indptr", "indices", "counts", "measure"))
    poff, sums = gpu.fetch_batch(br)
    return out, poff, sums, r.sum, r.nout, r.nonempty


def _same(x, y):
    (o1, p1, s1, t1, n1, e1), (o2, p2, s2, t2, n2, e2) = x, y
    for k in o1:
        assert np.array_equal(o1[k], o2[k]), k
    assert np.array_equal(p1, p2)
    assert np.array_equal(s1, s2)
    assert (t1, n1, e1) == (t2, n2, e2)


@pytest.mark.parametrize("dim", [2, 3])
def test_replays_match_the_eager_count_pass(dim):
    ba, bb = _batch(dim, seed=5)
    other = (gpu.DeviceMesh(*meshgen.split_grid(dim, 3, 9, mesh_id=0, diagonal=0)),
             gpu.DeviceMesh(*meshgen.split_grid(dim, 2, 9, mesh_id=1, diagonal=1)))
    ctx = gpu.Context()
    ref = _outputs(*ctx.supermesh_batched(ba, bb, flags=gpu.SORTED), dim)
    assert ctx.count_shape(dim)["block"] == 0    # nothing measured before the first GRAPH call
    for rep in range(4):
        r, br = ctx.supermesh_batched(ba, bb, flags=gpu.SORTED | gpu.GRAPH)
        print(f"[graph] {dim}D replay {rep}: {r.npairs} pairs, recorded graph {r.count_graph}", flush=True)
        assert r.count_graph == 1
        _same(_outputs(r, br, dim), ref)
        # the caller's own eager work in between (a smaller call: the same buffers)
        ctx.supermesh(*other)
        gpu.synchronize()
    shape = ctx.count_shape(dim)
    # 2D grid 0: one block per `block` pairs
    assert shape["block"] in (64, 128, 256) and (shape["grid"] >= 8 or (dim == 2 and shape["grid"] == 0))
    assert shape["ms"] > 0
    assert shape["lds_depths"] == (0 if dim == 2 else shape["lds_depths"])
    if dim == 3:
        assert shape["lds_depths"] in (2, 3)


@pytest.mark.parametrize("dim", [2, 3])
def test_replay_follows_moving_meshes(dim):
    ba, bb = _batch(dim, seed=6)
    ctx = gpu.Context()
    r, _ = ctx.supermesh_batched(ba, bb, flags=gpu.SORTED | gpu.GRAPH)
    assert r.count_graph == 1
    pos = ba.positions.numpy()
    rng = np.random.default_rng(7)
    for step in range(3):
        # shift every A mesh by a small rigid offset (the unit-square / cube
        # meshes stay valid) and rewrite the positions in place
        moved = pos + rng.uniform(-0.01, 0.01, size=(1, dim))
        ba.positions.write(moved)
        got = _outputs(*ctx.supermesh_batched(ba, bb, flags=gpu.SORTED | gpu.GRAPH), dim)
        ref = _outputs(*ctx.supermesh_batched(ba, bb, flags=gpu.SORTED), dim)
        print(f"[graph] {dim}D moved mesh, step {step}: {got[0]['counts'].size} pairs", flush=True)
        _same(got, ref)


def test_graph_refuses_other_passes():
    pa, ea = meshgen.split_grid(3, 2, 1)
    A = gpu.DeviceMesh(pa, ea)
    ctx = gpu.Context()
    for extra in (gpu.ELEMENTS, gpu.MASS):
        with pytest.raises(gpu.GpuError):
            ctx.supermesh(A, A, flags=gpu.GRAPH | extra)
