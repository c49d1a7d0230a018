"""The learned output-size predictor: MFMA inference and training kernels
against a numpy restatement, and the single-pass (LEARNED) tet clip against
the plain count-then-fill path -- every pair's output elements identical
whatever the weights, in the run layout (runs placed apart) and compacted
(LIBSUPERMESH_GPU_DENSE, libsupermesh_gpu_compact: the plain arrays byte for
byte)."""
import numpy as np
import pytest

import mlp_ref as R
import oracle_lib as O
from libsupermesh_amd import gpu, meshgen

pytestmark = pytest.mark.gpu


def runs_end_to_end(out):
    """Every pair's run (elements[offsets[p] .. offsets[p] + counts[p])), in
    pair order: the plain layout's element array."""
    offs, cnt, el = out["offsets"], out["counts"].astype(np.int64), out["elements"]
    excl = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    idx = np.repeat(offs[:-1] - excl, cnt) + np.arange(int(cnt.sum()))
    return el[idx]


def check_runs_apart(r, out):
    """Runs lie inside the element array and do not overlap."""
    offs, cnt = out["offsets"][:-1], out["counts"]
    nz = cnt > 0
    start, end = offs[nz], offs[nz] + cnt[nz]
    assert np.all(start >= 0) and np.all(end <= r.elements_size)
    order = np.argsort(start, kind="stable")
    assert np.all(end[order][:-1] <= start[order][1:])


@pytest.fixture(scope="module")
def meshes():
    pa, ea = meshgen.split_grid(3, 6, 21, mesh_id=0, diagonal=0)
    pb, eb = meshgen.split_grid(3, 5, 21, mesh_id=1, diagonal=1)
    return pa, ea, pb, eb


def test_inference_matches_numpy(meshes):
    pa, ea, pb, eb = meshes
    ip, ix = O.brute_force(pa, ea, pb, eb)
    rows, cols = O.csr_to_pairs(ip, ix)
    ctx = gpu.Context()
    params = R.init(3, scale=3.0)
    ctx.predictor_set(params)
    ctx.predictor_quantile(0.99)
    got = ctx.predictor_infer(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb), rows, cols)
    x = R.features(pa[ea[rows]], pb[eb[cols]])
    logits, _ = R.forward(params, x)
    ref = R.budgets(logits)
    assert len(set(ref.tolist())) > 3   # the weights give varied budgets
    agree = np.mean(got == ref)
    assert agree >= 0.97, agree          # off only where BF16 rounding flips a threshold
    assert np.all(np.abs(got - ref)[got != ref] <= 81)
    assert np.array_equal(ctx.predictor_get(), params)


def test_training_step_matches_numpy_gradient():
    rng = np.random.default_rng(4)
    n = 4096
    x = rng.uniform(0, 1, (n, 24)).astype(np.float32)
    y = rng.integers(0, 33, n).astype(np.int32)
    params = R.init(5)
    ctx = gpu.Context()
    ctx.predictor_set(params)
    ctx.predictor_add_samples(x, y)
    loss = ctx.predictor_train(epochs=1, batch=n, lr=1e-3, seed=0)   # one Adam step
    ref_loss, grad = R.loss_and_grads(params, R.bf16(x), y)
    assert abs(loss[0] - ref_loss) <= 1e-3 * ref_loss
    step = ctx.predictor_get() - params
    # Adam's first step moves each weight by lr * sign(gradient)
    big = np.abs(grad) > 1e-2 * np.abs(grad).max()
    assert big.sum() > 1000
    agree = np.mean(np.sign(step[big]) == -np.sign(grad[big]))
    assert agree >= 0.99, agree
    assert np.allclose(np.abs(step[big]), 1e-3, rtol=1e-2)


def test_training_learns_clip_counts(meshes):
    pa, ea, pb, eb = meshes
    ctx = gpu.Context()
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    n = ctx.predictor_collect(A, B, 1 << 20)
    assert n > 2000
    loss = ctx.predictor_train(epochs=12, batch=512, lr=3e-3, seed=7)
    assert np.all(np.isfinite(loss))
    assert loss[-1] < loss[len(loss) // 2] < loss[0], loss
    assert loss[-1] < 0.8 * loss[0], loss


@pytest.mark.parametrize("weights", ["trained", "random", "zero"])
@pytest.mark.parametrize("quantile", [0.5, 0.99])
@pytest.mark.parametrize("arena", ["default", "spill", "none", "tiny"])
def test_learned_single_pass_is_identical(meshes, weights, quantile, arena):
    pa, ea, pb, eb = meshes
    ctx = gpu.Context()
    if arena == "spill":
        ctx.predictor_arena(0.0, 1 << 20)   # a few arena tets per lane: some pairs are clipped again
    elif arena == "none":
        ctx.predictor_arena(0.0, 0)   # no arena: every pair beyond its budget is clipped again
    elif arena == "tiny":
        ctx.predictor_arena(0.0, 1)
        ctx.predictor_spill(0.0, 1)   # a one-tet spill reserve: the plain fill of every pair
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    if weights == "trained":
        ctx.predictor_collect(A, B, 1 << 20)
        ctx.predictor_train(epochs=6, batch=512, lr=3e-3, seed=1)
    elif weights == "random":
        ctx.predictor_set(R.init(9, scale=3.0))
    else:
        ctx.predictor_set(np.zeros(gpu.PREDICTOR_PARAMS, np.float32))
    ctx.predictor_quantile(quantile)
    what = ["indptr", "indices", "counts", "measure", "offsets", "elements"]
    r0 = ctx.supermesh(A, B, flags=gpu.ELEMENTS | gpu.SORTED)
    plain = gpu.fetch(r0, 3, what)
    assert r0.elements_size == r0.nout
    n0 = (r0.npairs, r0.nout, r0.sum)
    r1 = ctx.supermesh(A, B, flags=gpu.ELEMENTS | gpu.SORTED | gpu.LEARNED)
    learned = gpu.fetch(r1, 3, what)
    assert (r1.npairs, r1.nout, r1.sum) == n0
    for k in ("indptr", "indices", "counts", "measure"):
        assert np.array_equal(plain[k], learned[k]), k
    refilled = r1.elements_size == r1.nout and np.array_equal(learned["offsets"], plain["offsets"])
    if not refilled:
        check_runs_apart(r1, learned)
    assert np.array_equal(runs_end_to_end(learned), plain["elements"])
    # compacted after the call, and in the call: the plain arrays byte for byte
    ctx.compact(r1)
    assert (r1.elements_size, r1.nout) == (r0.nout, r0.nout)
    compact = gpu.fetch(r1, 3, ["offsets", "elements"])
    r2 = ctx.supermesh(A, B, flags=gpu.ELEMENTS | gpu.SORTED | gpu.LEARNED | gpu.DENSE)
    dense = gpu.fetch(r2, 3, ["offsets", "elements"])
    for k in ("offsets", "elements"):
        assert np.array_equal(compact[k], plain[k]), k
        assert np.array_equal(dense[k], plain[k]), k
    if weights == "zero" and quantile == 0.99:
        # uniform prediction: the 0.99 quantile is the last class, budget = maximum
        assert r1.learned_overflow == 0 and not refilled
    if weights == "random":
        assert r1.learned_overflow > 0    # outputs beyond their budgets: joined or re-clipped runs
        assert refilled == (arena == "tiny")
    assert r1.learned_slots >= 0
    assert 0.0 < r1.learned_quantile <= quantile + 1e-6   # a float32 level


@pytest.mark.parametrize("case,route", [("roomy", 1), ("no_budgets", 2), ("tiny_spill", 3)])
def test_dense_routes(meshes, case, route):
    """LIBSUPERMESH_GPU_DENSE's three ways to the plain layout, each forced,
    each byte for byte the plain arrays: the runs moved within their buffer
    (every budget the maximum: no pair's plain start beyond its run's), copied
    to a second buffer (every budget 0: the plain layout lies beyond the
    runs' start), every pair filled again (a one-tet spill reserve that the
    pairs clipped again outgrow)."""
    pa, ea, pb, eb = meshes
    ctx = gpu.Context()
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    if case == "tiny_spill":
        ctx.predictor_set(R.init(9, scale=3.0))
        ctx.predictor_arena(0.0, 0)
        ctx.predictor_spill(0.0, 1)
        ctx.predictor_quantile(0.5)
    else:
        ctx.predictor_set(np.zeros(gpu.PREDICTOR_PARAMS, np.float32))
        ctx.predictor_quantile(0.99 if case == "roomy" else 0.01)
    what = ["counts", "offsets", "elements"]
    plain = gpu.fetch(ctx.supermesh(A, B, flags=gpu.ELEMENTS), 3, what)
    r = ctx.supermesh(A, B, flags=gpu.ELEMENTS | gpu.LEARNED | gpu.DENSE)
    print(f"[dense] {case}: route {r.dense_route}, {r.learned_slots} slots, {r.learned_overflow} overflow, "
          f"{r.nout} tets", flush=True)
    assert r.dense_route == route
    assert r.elements_size == r.nout
    got = gpu.fetch(r, 3, what)
    for k in what:
        assert np.array_equal(got[k], plain[k]), k
    if case == "no_budgets":
        assert r.learned_slots == 0 and r.learned_overflow > 0
    ctx.compact(r)   # already plain: left as it is
    assert r.elements_size == r.nout and np.array_equal(gpu.fetch(r, 3, ["elements"])["elements"], plain["elements"])


@pytest.mark.parametrize("weights", ["trained", "random"])
def test_learned_single_pass_with_mass_is_identical(meshes, weights):
    """LEARNED with MASS (the single pass that also integrates the mixed mass
    blocks): counts, elements, blocks and the assembled mass CSR equal the
    plain path's."""
    pa, ea, pb, eb = meshes
    ctx = gpu.Context()
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    if weights == "trained":
        ctx.predictor_collect(A, B, 1 << 20)
        ctx.predictor_train(epochs=6, batch=512, lr=3e-3, seed=1)
    else:
        ctx.predictor_set(R.init(9, scale=3.0))
    ctx.predictor_arena(0.0, 0)
    ctx.predictor_spill(0.0, 1)   # force overflow pairs through the fallback too
    what = ["indptr", "indices", "counts", "measure", "offsets", "elements", "blocks"]
    flags = gpu.ELEMENTS | gpu.SORTED | gpu.MASS | gpu.BLOCKS
    r0 = ctx.supermesh(A, B, flags=flags)
    plain = gpu.fetch(r0, 3, what)
    m0 = gpu.fetch_mass(r0, pa.shape[0])
    r1 = ctx.supermesh(A, B, flags=flags | gpu.LEARNED | gpu.DENSE)
    learned = gpu.fetch(r1, 3, what)
    m1 = gpu.fetch_mass(r1, pa.shape[0])
    assert r1.learned_quantile > 0
    for k in what:
        assert np.array_equal(plain[k], learned[k]), k
    for x, y in zip(m0, m1):
        assert np.array_equal(x, y)
    # and the blocks against the oracle's quadrature
    rows, cols = O.csr_to_pairs(plain["indptr"], plain["indices"])
    ref = O.mass_pairs(pa, ea, pb, eb, rows, cols)
    scale = np.abs(ref).max()
    assert np.abs(learned["blocks"] - ref).max() <= 1e-12 * scale


@pytest.mark.parametrize("dense", [False, True])
def test_learned_batched_is_identical(dense):
    """LIBSUPERMESH_GPU_LEARNED on a batch of 3D mesh pairs (one batched call,
    the walk over every mesh's pairs at once): each mesh's counts and output
    tets equal those of one plain count-then-fill call per mesh pair, in the
    run layout (runs read where their offsets put them) and compacted."""
    pairs = [p for p in meshgen.batched_pairs(count=24, seed=5, emin=50, emax=3000) if p[0] == 3]
    assert len(pairs) >= 6
    ctx = gpu.Context()
    ctx.predictor_set(R.init(9, scale=3.0))   # budgets that many runs outgrow
    ctx.predictor_quantile(0.5)
    BA = gpu.DeviceMeshBatch([(p[1], p[2]) for p in pairs])
    BB = gpu.DeviceMeshBatch([(p[3], p[4]) for p in pairs])
    flags = gpu.ELEMENTS | gpu.SORTED | gpu.LEARNED | (gpu.DENSE if dense else 0)
    r, br = ctx.supermesh_batched(BA, BB, flags=flags)
    poff, sums = gpu.fetch_batch(br)
    got = gpu.fetch(r, 3, ["counts", "offsets", "elements"])
    assert r.learned_overflow > 0 and r.learned_quantile > 0
    if dense:
        assert r.elements_size == r.nout and r.dense_route in (1, 2, 3)
    else:
        check_runs_apart(r, got)
    for m, (_, pa, ea, pb, eb) in enumerate(pairs):
        rp = ctx.supermesh(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb), flags=gpu.ELEMENTS | gpu.SORTED)
        plain = gpu.fetch(rp, 3, ["counts", "elements"])
        p0, p1 = int(poff[m]), int(poff[m + 1])
        assert p1 - p0 == rp.npairs, m
        cnt = got["counts"][p0:p1]
        assert np.array_equal(cnt, plain["counts"]), m
        runs = [got["elements"][got["offsets"][p]:got["offsets"][p] + got["counts"][p]] for p in range(p0, p1)]
        mine = np.concatenate(runs) if runs else np.zeros((0, 4, 3))
        assert np.array_equal(mine.reshape(-1, 4, 3), plain["elements"].reshape(-1, 4, 3)), m
        assert abs(sums[m] - rp.sum) <= 1e-12, m
