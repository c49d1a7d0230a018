"""Benchmark: clipped candidate pairs per second on one MI355X.

A step is one full supermesh pass over resident meshes: element staging,
hierarchy build, candidate search, FP64 clipping of every candidate pair,
reductions and (2D) P1 x P1 mixed-mass assembly -- one call of
libsupermesh_gpu_supermesh.  The headline workload is the 2D split-grid pair
1024^2 (2,097,152 triangles) vs 1000^2 (2,000,000 triangles), jitter 0.3h,
seed 1; the 3D pair 64^3 x 6 (1,572,864 tets) vs 60^3 x 6 (1,296,000 tets),
jitter 0.25h, seed 2, is reported beside it.

Prints one JSON line.  `value` = candidate pairs clipped per second,
end to end (inputs resident in HBM when timing starts).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "clipped candidate pairs/s (2D tri, 3D tet); learned+batched vs plain, 1 GPU"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix), FMA counted as 2 flops
HBM_PEAK_GBS = 8000.0

# FP64 arithmetic operations (add, sub, mul, div, sqrt; compares, abs, min
# and max are not counted) the reference algorithm executes per candidate
# pair, counted with the instrumented CPU restatement (oracle/count_flops.py)
# on these workloads: clip + (2D) the mixed-mass quadrature.
FLOPS_PER_PAIR = {"tri_1024_vs_1000": 108.6 + 268.0, "tet_64_vs_60": 544.6}

# Algorithmic bytes per candidate pair of the clip count pass: the B element
# gathered (tri 6 doubles + 3 ids = 60 B; tet 12 doubles + 4 ids = 112 B),
# the pair's CSR index (4 B), its count (4 B) and measure (8 B); with the
# mixed mass (2D) also its 3x3 block (72 B).
BYTES_PER_PAIR = {"tri_1024_vs_1000": 60 + 4 + 4 + 8 + 72, "tet_64_vs_60": 112 + 4 + 4 + 8}

# HBM bytes per launch of the same kernel on the same workload, from the two
# counter runs summarised in profiles/traffic_tri1024_tet64.md (FETCH_SIZE x
# 1.9992 + WRITE_SIZE x 1.0, the factors a 256 MiB copy calibrates).
PROFILED_TRAFFIC = {"tri_1024_vs_1000": 3.55e9, "tet_64_vs_60": 28.46e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="tri_1024_vs_1000")
    p.add_argument("--no-3d", action="store_true", help="skip the 3D workload")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-dropin", action="store_true", help="skip the drop-in host-array finder timing")
    p.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    p.add_argument("--no-mixed", action="store_true", help="skip the tri/quad and tet/hex workloads")
    p.add_argument("--batch-count", type=int, default=4096, help="mesh pairs in the batched workload")
    p.add_argument("--no-learned", action="store_true", help="skip the learned single-pass workload")
    p.add_argument("--per-call-only", action="store_true",
                   help="only time one plain call per small mesh pair (first 256 of the batched generator)")
    p.add_argument("--learned-workload", default="tet_96_vs_90")
    p.add_argument("--weights", default=os.path.join(ROOT, "libsupermesh_amd", "data", "tet_counts_mlp.f32"))
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a plain device copy of known size (PMC calibration)")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# A one-GPU box gives its process a share of 16 host cores; worker pools are
# sized to that share whatever the machine's core count.
BOX_CPU_SHARE = 16


def host_info():
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": aff, "cpu_model": model,
            "threads_used": max(1, min(aff, BOX_CPU_SHARE))}


def host_threads():
    return host_info()["threads_used"]


def gpu_workload(ctx, name, steps, warmup, mass):
    from libsupermesh_amd import gpu, meshgen
    t0 = time.time()
    pa, ea, pb, eb = meshgen.config_meshes(name)
    log(f"[bench] {name}: meshes {ea.shape[0]} x {eb.shape[0]} elements ({time.time() - t0:.1f}s)")
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    flags = gpu.TIMING | (gpu.MASS if mass else 0)
    for _ in range(warmup):
        r = ctx.supermesh(A, B, flags=flags)
    gpu.synchronize()
    t0 = time.perf_counter()
    clip_ms = []
    phases = []
    for _ in range(steps):
        r = ctx.supermesh(A, B, flags=flags)
        clip_ms.append(r.ms[7])
        phases.append(list(r.ms))
    gpu.synchronize()
    dt = (time.perf_counter() - t0) / steps
    err = abs(r.sum - 1.0)
    log(f"[bench] {name}: {r.npairs} pairs, {r.nonempty} non-empty, {r.nout} elements, "
        f"{dt * 1e3:.2f} ms/step, |sum-1|={err:.2e}")
    names = ["stage", "hierarchy", "search", "clip", "mass", None, "total", "clip_kernel"]
    avg = [sum(p[i] for p in phases) / steps for i in range(8)]
    return {
        "workload": name, "elements_a": int(ea.shape[0]), "elements_b": int(eb.shape[0]),
        "pairs": int(r.npairs), "nonempty": int(r.nonempty), "supermesh_elements": int(r.nout),
        "max_per_pair": int(r.max_count), "sum_abs_err": err, "ms_per_step": dt * 1e3,
        "pairs_per_s": r.npairs / dt, "elements_per_s": r.nout / dt,
        "clip_kernel_ms": sum(clip_ms) / len(clip_ms),
        "phase_ms": {n: round(v, 3) for n, v in zip(names, avg) if n},
        "mass_nnz": int(r.mass_nnz) if mass else None,
        "_meshes": (pa, ea, pb, eb),
    }


MIXED_WORKLOADS = {
    # name: (dim, simplex grid n, cell grid n)
    "tri_1024_vs_quad_1000": (2, 1024, 1000),
    "tet_64_vs_hex_60": (3, 64, 60),
}


def mixed_workload(ctx, steps, warmup):
    """Simplex mesh A against a quadrilateral / hexahedral mesh B, every
    candidate pair clipped as intersect_elements does (tri x quad, tet x hex),
    counts, measures and the compensated sum, R-tree (closed) candidates."""
    from libsupermesh_amd import gpu, meshgen
    out = {}
    for name, (dim, na, nb) in MIXED_WORKLOADS.items():
        pa, ea = meshgen.split_grid(dim, na, 6, mesh_id=0)
        pb, eb = meshgen.cell_grid(dim, nb, 6, mesh_id=1)
        A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
        flags = gpu.TIMING
        for _ in range(warmup):
            r = ctx.supermesh_mixed(A, B, flags=flags)
        gpu.synchronize()
        t0 = time.perf_counter()
        phases = []
        for _ in range(steps):
            r = ctx.supermesh_mixed(A, B, flags=flags)
            phases.append(list(r.ms))
        gpu.synchronize()
        dt = (time.perf_counter() - t0) / steps
        err = abs(r.sum - 1.0)
        log(f"[bench] {name}: {r.npairs} pairs, {r.nout} elements, {dt * 1e3:.2f} ms/step, |sum-1|={err:.2e}")
        names = ["stage", "hierarchy", "search", "clip"]
        out[name] = {"elements_a": int(ea.shape[0]), "elements_b": int(eb.shape[0]), "pairs": int(r.npairs),
                     "nonempty": int(r.nonempty), "supermesh_elements": int(r.nout), "sum_abs_err": err,
                     "ms_per_step": dt * 1e3, "pairs_per_s": r.npairs / dt,
                     "phase_ms": {n: round(sum(p[i] for p in phases) / steps, 3) for i, n in enumerate(names)}}
    return out


def batched_workload(ctx, steps, warmup, count=4096):
    """The batched workload: `count` independent mesh pairs (2D or 3D with
    equal odds, element counts log-uniform in [50, 50000] per mesh, seed 3).
    Plain = one libsupermesh_gpu_supermesh call per mesh pair; batched = one
    libsupermesh_gpu_supermesh_batched call per dimension.  Both compute the
    candidate pairs, their clips and per-pair / per-mesh measure sums."""
    import numpy as np
    from libsupermesh_amd import gpu, meshgen
    t0 = time.time()
    pairs = meshgen.batched_pairs(count=count, seed=3)
    log(f"[bench] batched: {len(pairs)} mesh pairs generated ({time.time() - t0:.1f}s)")
    plain = [(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)) for _, pa, ea, pb, eb in pairs]
    batches = {}
    for dim in (2, 3):
        sel = [k for k, p in enumerate(pairs) if p[0] == dim]
        if sel:
            batches[dim] = (sel, gpu.DeviceMeshBatch([(pairs[k][1], pairs[k][2]) for k in sel]),
                            gpu.DeviceMeshBatch([(pairs[k][3], pairs[k][4]) for k in sel]))
    log(f"[bench] batched: meshes resident ({time.time() - t0:.1f}s)")

    def run_plain():
        tot, sums = 0, np.zeros(len(pairs))
        for k, (A, B) in enumerate(plain):
            r = ctx.supermesh(A, B)
            tot += r.npairs
            sums[k] = r.sum
        return tot, sums

    def run_batched(flags=0):
        tot, sums = 0, np.zeros(len(pairs))
        for dim, (sel, BA, BB) in batches.items():
            r, br = ctx.supermesh_batched(BA, BB, flags=flags)
            tot += r.npairs
            sums[sel] = gpu.fetch_batch(br)[1]
            if flags & gpu.GRAPH:
                graphs[dim] = int(r.count_graph)
        return tot, sums

    graphs = {}
    out = {}
    # batched_graph: the count pass of each batched call as a recorded graph,
    # in the launch shape the first call measured fastest (measured in the
    # warm-up, outside the timed steps)
    for name, fn in (("plain", run_plain), ("batched", run_batched),
                     ("batched_graph", lambda: run_batched(gpu.GRAPH))):
        for _ in range(warmup):
            fn()
        gpu.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            npairs, sums = fn()
        gpu.synchronize()
        dt = (time.perf_counter() - t1) / steps
        out[name] = {"ms_per_step": dt * 1e3, "pairs": int(npairs), "pairs_per_s": npairs / dt,
                     "_sums": sums}
        log(f"[bench] batched workload, {name}: {npairs} pairs, {dt * 1e3:.1f} ms/step")
    dev = np.abs(out["plain"]["_sums"] - out["batched"]["_sums"])
    ref = np.maximum(1.0, np.abs(out["plain"]["_sums"]))
    unit_err = float(np.abs(out["batched"]["_sums"] - 1.0).max())   # every mesh covers [0,1]^d
    graph_same = bool(np.array_equal(out["batched"]["_sums"], out["batched_graph"]["_sums"]))
    for v in out.values():
        v.pop("_sums")
    n2 = len(batches.get(2, ([],))[0])
    return {
        "workload": f"{len(pairs)} mesh pairs ({n2} 2D, {len(pairs) - n2} 3D), "
                    "elements log-uniform in [50, 50000] per mesh, seed 3",
        "plain": out["plain"], "batched": out["batched"], "batched_graph": out["batched_graph"],
        "speedup": out["plain"]["ms_per_step"] / out["batched_graph"]["ms_per_step"],
        "speedup_without_graph": out["plain"]["ms_per_step"] / out["batched"]["ms_per_step"],
        "graph": {"count_pass_replayed": graphs, "same_sums_as_eager": graph_same,
                  "launch_shapes": {f"{d}d": ctx.count_shape(d) for d in batches}},
        "same_pairs": out["plain"]["pairs"] == out["batched"]["pairs"] == out["batched_graph"]["pairs"],
        "max_rel_sum_diff": float((dev / ref).max()),
        "max_abs_sum_err": unit_err,
    }


def learned_workload(ctx, name, steps, warmup, weights, quantiles=(0.7, 0.8, 0.85, 0.9, 0.99)):
    """Materialised tet supermesh (LIBSUPERMESH_GPU_ELEMENTS): plain
    count-then-fill against the single pass sized by the learned predictor
    (LIBSUPERMESH_GPU_LEARNED), which leaves each pair's run where its budget
    placed it, and against the same single pass compacted into the plain
    layout in the call (LIBSUPERMESH_GPU_DENSE).  Outputs are compared by
    device checksums of the per-pair counts, the offsets and every output
    tet, the single pass's runs after compaction."""
    import numpy as np
    from libsupermesh_amd import gpu, meshgen
    t0 = time.time()
    pa, ea, pb, eb = meshgen.config_meshes(name)
    log(f"[bench] learned {name}: meshes {ea.shape[0]} x {eb.shape[0]} ({time.time() - t0:.1f}s)")
    A, B = gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)
    ctx.predictor_set(np.fromfile(weights, dtype=np.float32))
    out = {}
    # arena: None = the default (a quarter of the expected output, spread over
    # the walk's lanes), 0 = none (every pair beyond its budget clipped again)
    modes = [("plain", gpu.ELEMENTS | gpu.TIMING, None, None)]
    modes += [(f"learned_q{q}", gpu.ELEMENTS | gpu.LEARNED | gpu.TIMING, q, None) for q in quantiles]
    modes += [("learned_q0.9_noarena", gpu.ELEMENTS | gpu.LEARNED | gpu.TIMING, 0.9, 0)]
    modes += [("learned_dense", gpu.ELEMENTS | gpu.LEARNED | gpu.DENSE | gpu.TIMING, 0.99, None)]
    modes += [("learned_dense_noarena", gpu.ELEMENTS | gpu.LEARNED | gpu.DENSE | gpu.TIMING, 0.9, 0)]

    def sums(r):
        return (ctx.checksum(r.counts, 4 * r.npairs), ctx.checksum(r.offsets, 8 * (r.npairs + 1)),
                ctx.checksum(r.elements, 96 * r.nout))

    for mode, flags, q, arena in modes:
        if q is not None:
            ctx.predictor_quantile(q)
            ctx.predictor_arena(*((0.25, 1 << 20) if arena is None else (0.0, 0)))
        for _ in range(warmup):
            ctx.supermesh(A, B, flags=flags)
        gpu.synchronize()
        t1 = time.perf_counter()
        walk = []
        pred = []
        phases = []
        for _ in range(steps):
            r = ctx.supermesh(A, B, flags=flags)
            walk.append(r.ms[7])
            pred.append(r.ms[5])
            phases.append(list(r.ms))
        gpu.synchronize()
        dt = (time.perf_counter() - t1) / steps
        if q is not None and not flags & gpu.DENSE:
            # the runs of this very call, compacted (untimed) for the comparison
            level = r.learned_quantile
            slots, ovf, size = int(r.learned_slots), int(r.learned_overflow), int(r.elements_size)
            r = ctx.supermesh(A, B, flags=flags | gpu.DENSE)
        out[mode] = {"ms_per_step": dt * 1e3, "pairs": int(r.npairs), "supermesh_elements": int(r.nout),
                     "pairs_per_s": r.npairs / dt, "elements_per_s": r.nout / dt,
                     "first_walk_ms": sum(walk) / steps, "_sums": sums(r),
                     "phase_ms": {k: round(sum(p[i] for p in phases) / steps, 3)
                                  for k, i in (("search", 2), ("clip", 3), ("total", 6))}}
        if q is not None:
            if flags & gpu.DENSE:
                level = r.learned_quantile
                slots, ovf, size = int(r.learned_slots), int(r.learned_overflow), int(r.elements_size)
                out[mode]["dense_route"] = {1: "moved in place", 2: "copied", 3: "filled again"}.get(
                    int(r.dense_route), "plain path")
            out[mode].update({"quantile_requested": q, "quantile_used": round(float(level), 6),
                              "predictor_ms": sum(pred) / steps, "slots": slots, "overflow_pairs": ovf,
                              "elements_array": size})
        log(f"[bench] learned {name}, {mode}: {r.npairs} pairs, {r.nout} tets, {dt * 1e3:.1f} ms/step, "
            f"phases {out[mode]['phase_ms']}")
    ref = out["plain"].pop("_sums")
    same = all(v.pop("_sums") == ref for k, v in out.items() if k != "plain")
    runs = [k for k in out if k.startswith("learned_q")]
    best = min(runs, key=lambda k: out[k]["ms_per_step"])
    dense = min((k for k in out if k.startswith("learned_dense")), key=lambda k: out[k]["ms_per_step"])
    return {"workload": name, "plain": out["plain"], "learned": out[best], "learned_mode": best,
            "speedup": out["plain"]["ms_per_step"] / out[best]["ms_per_step"],
            "learned_dense": out[dense], "learned_dense_mode": dense,
            "speedup_dense": out["plain"]["ms_per_step"] / out[dense]["ms_per_step"],
            "identical_outputs": same, "weights": os.path.relpath(weights, ROOT),
            "quantile_sweep": {k: {"ms_per_step": round(out[k]["ms_per_step"], 3),
                                   "quantile_used": out[k]["quantile_used"], "slots": out[k]["slots"],
                                   "overflow_pairs": out[k]["overflow_pairs"]}
                               for k in runs + [k for k in out if k.startswith("learned_dense")]}}


def clip_roofline(name, w):
    """Roofline of the clip count pass (the dominant kernel of the 2D step
    and of the 3D step), timed with HIP events on the library's stream over
    the timed steps.  Both its ceilings are given: HBM (the bound at this
    arithmetic intensity) and FP64."""
    secs = w["clip_kernel_ms"] / 1e3
    bpp, fpp = BYTES_PER_PAIR.get(name), FLOPS_PER_PAIR.get(name)
    if not bpp or secs <= 0:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel_ms": w["clip_kernel_ms"]}
    gbs = bpp * w["pairs"] / secs / 1e9
    out = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "traffic": PROFILED_TRAFFIC.get(name), "kernel": "k_clip_tri count pass with mixed mass"
           if name.startswith("tri") else "k_clip_tet_walk count pass",
           "bytes_per_pair": bpp, "pairs_per_launch": w["pairs"], "kernel_ms": w["clip_kernel_ms"]}
    if fpp:
        tf = fpp * w["pairs"] / secs / 1e12
        # bit-exactness forbids fused multiply-adds: half the FMA-counted peak
        # is the attainable one for this code
        out["fp64"] = {"achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS,
                       "peak_without_fma": FP64_PEAK_TFLOPS / 2, "frac_without_fma": tf / (FP64_PEAK_TFLOPS / 2),
                       "flops_per_pair": fpp}
    return out


def learned_batched_workload(ctx, steps, warmup, weights, count=1024):
    """The materialised supermesh (ELEMENTS, rows sorted) of `count` mesh pairs
    of the batched generator: plain = one count-then-fill call per mesh pair;
    learned+batched = one batched call per dimension, the 3D one sized by the
    predictor (its runs left where the budgets placed them).  Outputs are
    compared mesh by mesh by device checksums of each mesh's per-pair counts
    and output elements, the single pass compacted in the call."""
    import numpy as np
    from libsupermesh_amd import gpu, meshgen
    t0 = time.time()
    pairs = meshgen.batched_pairs(count=count, seed=3)
    plain = [(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb)) for _, pa, ea, pb, eb in pairs]
    batches = {}
    for dim in (2, 3):
        sel = [k for k, p in enumerate(pairs) if p[0] == dim]
        if sel:
            batches[dim] = (sel, gpu.DeviceMeshBatch([(pairs[k][1], pairs[k][2]) for k in sel]),
                            gpu.DeviceMeshBatch([(pairs[k][3], pairs[k][4]) for k in sel]))
    ctx.predictor_set(np.fromfile(weights, dtype=np.float32))
    log(f"[bench] learned+batched: {len(pairs)} mesh pairs resident ({time.time() - t0:.1f}s)")
    base = gpu.ELEMENTS | gpu.SORTED

    def sums_plain():
        out = []
        for A, B in plain:
            r = ctx.supermesh(A, B, flags=base)
            dim = A.dim
            out.append((r.npairs, r.nout, ctx.checksum(r.counts, 4 * r.npairs),
                        ctx.checksum(r.elements, 8 * (dim + 1) * dim * r.nout)))
        return out

    def sums_batched():
        out = [None] * len(pairs)
        for dim, (sel, BA, BB) in batches.items():
            # the single pass's runs compacted in the call, for the comparison
            r, br = ctx.supermesh_batched(BA, BB, flags=base | (gpu.LEARNED | gpu.DENSE if dim == 3 else 0))
            poff, _ = gpu.fetch_batch(br)
            offs = gpu.device_to_numpy(r.offsets, r.npairs + 1, np.int64)
            esz = 8 * (dim + 1) * dim
            for m, k in enumerate(sel):
                p0, p1 = int(poff[m]), int(poff[m + 1])
                e0, e1 = int(offs[p0]), int(offs[p1])
                out[k] = (p1 - p0, e1 - e0, ctx.checksum((r.counts or 0) + 4 * p0, 4 * (p1 - p0)),
                          ctx.checksum((r.elements or 0) + esz * e0, esz * (e1 - e0)))
        return out

    def run_plain():
        tot = nout = 0
        for A, B in plain:
            r = ctx.supermesh(A, B, flags=base)
            tot += r.npairs
            nout += r.nout
        return tot, nout

    def run_batched():
        tot = nout = 0
        for dim, (sel, BA, BB) in batches.items():
            r, _ = ctx.supermesh_batched(BA, BB, flags=base | (gpu.LEARNED if dim == 3 else 0))
            tot += r.npairs
            nout += r.nout
        return tot, nout

    same = sums_plain() == sums_batched()
    out = {}
    for name, fn in (("plain", run_plain), ("learned_batched", run_batched)):
        for _ in range(warmup):
            fn()
        gpu.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            npairs, nout = fn()
        gpu.synchronize()
        dt = (time.perf_counter() - t1) / steps
        out[name] = {"ms_per_step": dt * 1e3, "pairs": int(npairs), "supermesh_elements": int(nout),
                     "pairs_per_s": npairs / dt, "elements_per_s": nout / dt}
        log(f"[bench] learned+batched workload, {name}: {npairs} pairs, {nout} elements, {dt * 1e3:.1f} ms/step")
    return {"workload": f"{len(pairs)} mesh pairs of the batched generator, materialised supermesh, rows sorted",
            "plain": out["plain"], "learned_batched": out["learned_batched"],
            "speedup": out["plain"]["ms_per_step"] / out["learned_batched"]["ms_per_step"],
            "identical_outputs": same}


def cpu_baseline(name, meshes, mass):
    """The CPU restatement of the reference's serial loop on the host cores
    (tree finder, then the per-pair clip, and with mass the per-pair P1 x P1
    block of the reference tests' quadrature contracted with linear fields),
    one thread per core over contiguous ranges of A, on the whole workload;
    and one thread on a bounded sample (the first tenth of A's elements)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    pa, ea, pb, eb = meshes
    info = host_info()
    n = info["threads_used"]
    stats, s, (tb, tq) = O.supermesh(pa, ea, pb, eb, finder=0, nthreads=n, mass=mass)
    sub = ea[: max(1, ea.shape[0] // 10)]
    st1, _, (tb1, tq1) = O.supermesh(pa, sub, pb, eb, finder=0, nthreads=1, mass=mass)
    what = "tree query + clip" + (" + mixed-mass blocks (f^T M g)" if mass else "")
    out = {
        "value": stats["candidates"] / tq, "unit": "pairs/s", "cores": n, "kind": "port",
        "sample": (f"whole {name} workload: serial tree build on B ({tb:.2f} s, excluded), then {what} "
                   f"of all {stats['candidates']} candidate pairs on {n} threads ({tq:.2f} s); "
                   "single core: the same on the first tenth of A's elements "
                   f"({st1['candidates']} pairs, {tq1:.2f} s); no mass CSR assembly"),
        "end_to_end_pairs_per_s": stats["candidates"] / (tb + tq),
        "single_core_pairs_per_s": st1["candidates"] / tq1,
        "sum_abs_err": abs(s - 1.0),
        "host": info,
    }
    if mass:
        out["fmg_abs_err"] = abs(stats["fmg"] - 0.25)
    return out


def per_call_workload(ctx, count=256):
    """One plain libsupermesh_gpu_supermesh call per small mesh pair (the
    first `count` mesh pairs of the batched generator, seed 3: element counts
    log-uniform in [50, 50000], resident in HBM), timed per call with the
    library's phase times: the cost a caller pays who keeps the reference
    manual's loop of one call per mesh pair."""
    import numpy as np
    from libsupermesh_amd import gpu, meshgen
    pairs = meshgen.batched_pairs(count=count, seed=3)
    plain = [(d, gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb), ea.shape[0], eb.shape[0])
             for d, pa, ea, pb, eb in pairs]
    for _, A, B, _, _ in plain:
        ctx.supermesh(A, B)
    gpu.synchronize()
    out = {"workload": f"first {count} mesh pairs of the batched generator, one call each"}
    for dim in (2, 3):
        sel = [p for p in plain if p[0] == dim]
        if not sel:
            continue
        ph = np.zeros(8)
        t0 = time.perf_counter()
        for _, A, B, _, _ in sel:
            r = ctx.supermesh(A, B, flags=gpu.TIMING)
            ph += np.array(list(r.ms)[:8])
        gpu.synchronize()
        dt = (time.perf_counter() - t0) / len(sel)
        ph /= len(sel)
        out[f"{dim}d"] = {"calls": len(sel), "ms_per_call": dt * 1e3,
                          "mean_elements": [float(np.mean([p[3] for p in sel])), float(np.mean([p[4] for p in sel]))],
                          "phase_ms": {k: round(float(ph[i]), 3) for k, i in
                                       (("stage", 0), ("hierarchy", 1), ("search", 2), ("clip", 3),
                                        ("clip_kernel", 7))}}
        log(f"[bench] per call {dim}D: {len(sel)} calls, {dt * 1e3:.3f} ms/call, phases {out[f'{dim}d']['phase_ms']}")
    return out


def dropin_finder(meshes, steps=3):
    """The reference's drop-in C finder on host arrays:
    libsupermesh_tree_intersection_finder_set_input -> query_output ->
    get_output (C_bindings.F90:346-391), i.e. the H2D copy of both meshes, the
    device search and the D2H copy of the CSR, all inside the timed call."""
    import numpy as np
    from libsupermesh_amd import gpu
    pa, ea, pb, eb = meshes
    ea64, eb64 = ea.astype(np.int64), eb.astype(np.int64)
    ip, ix = gpu.tree_intersection_finder(pa, ea64, pb, eb64)
    t0 = time.perf_counter()
    for _ in range(steps):
        ip, ix = gpu.tree_intersection_finder(pa, ea64, pb, eb64)
    dt = (time.perf_counter() - t0) / steps
    return {"ms_per_call": dt * 1e3, "pairs": int(ix.shape[0]), "pairs_per_s": ix.shape[0] / dt,
            "note": "host arrays in, host CSR out (PCIe included); not the bench value"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on exactly one GPU: sharding A by Morton ranges "
                          "across GPUs is designed but not built; run with --gpus 1"}))
        sys.exit(2)
    import numpy as np
    from libsupermesh_amd import gpu
    ctx = gpu.Context(0)
    if args.calibrate_copy:
        n = 1 << 28  # 256 MiB device-to-device copy of known size
        src = gpu.DeviceArray(np.zeros(n, dtype=np.uint8))
        dst = gpu.DeviceArray(np.zeros(n, dtype=np.uint8))
        gpu._check(gpu.lib().libsupermesh_gpu_memcpy(dst.ptr, src.ptr, n, 3))
        gpu.synchronize()
        log(f"[bench] calibration copy of {n} bytes done")
    if args.per_call_only:
        print(json.dumps(per_call_workload(ctx)))
        ctx.close()
        return
    main_w = gpu_workload(ctx, args.workload, args.steps, args.warmup, mass=args.workload.startswith("tri"))
    extra = None
    if not args.no_3d and args.workload != "tet_64_vs_60":
        extra = gpu_workload(ctx, "tet_64_vs_60", max(1, args.steps), max(1, args.warmup), mass=False)
    mixed = None
    if not args.no_mixed:
        mixed = mixed_workload(ctx, max(1, min(args.steps, 3)), 1)
    batched = None
    if not args.no_batched:
        batched = batched_workload(ctx, max(1, min(args.steps, 3)), 1, count=args.batch_count)
    learned_batched = None
    if not args.no_learned and not args.no_batched:
        learned_batched = learned_batched_workload(ctx, max(1, min(args.steps, 2)), 1, args.weights,
                                                   count=max(1, args.batch_count // 4))
    per_call = None if args.no_batched else per_call_workload(ctx)
    learned = None
    if not args.no_learned:
        if not os.path.exists(args.weights):
            raise SystemExit(f"predictor weights not found: {args.weights}")
        learned = learned_workload(ctx, args.learned_workload, max(1, min(args.steps, 3)), 1, args.weights)
    roofline = clip_roofline(args.workload, main_w)
    if extra:
        extra["roofline"] = clip_roofline(extra["workload"], extra)
    for w in (main_w, extra):
        if w and not args.no_dropin:
            w["dropin_finder"] = dropin_finder(w["_meshes"])
            log(f"[bench] drop-in host finder {w['workload']}: {w['dropin_finder']['ms_per_call']:.1f} ms/call")
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(args.workload, main_w["_meshes"], mass=main_w["mass_nnz"] is not None)
        log(f"[bench] CPU baseline {cpu['value']:.3e} pairs/s on {cpu['cores']} threads")
        if extra:
            extra["cpu_baseline"] = cpu_baseline(extra["workload"], extra["_meshes"], mass=False)
            log(f"[bench] CPU baseline {extra['workload']}: {extra['cpu_baseline']['value']:.3e} pairs/s")
    for w in (main_w, extra):
        if w:
            w.pop("_meshes", None)
    out = {
        "metric": METRIC, "value": main_w["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": main_w["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "finder": "tree (strict AABB)",
                   "mixed_mass": main_w["mass_nnz"] is not None,
                   "pairs": main_w["pairs"], "supermesh_elements": main_w["supermesh_elements"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": main_w,
        "tet_64_vs_60": extra,
        "mixed_types": mixed,
        "batched": batched,
        "per_call": per_call,
        "learned": learned,
        "learned_batched": learned_batched,
    }
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
