"""Trains the learned output-size predictor and writes its weights.

Training data: the pairs the plain method clips fully (after the exact
separation test) on the 3D split-grid generator, 64^3 x 6 vs 60^3 x 6 tets,
jitter 0.25h, at seeds other than the benchmark's (default 10 and 11),
subsampled to --samples pairs in total; label = the pair's output count.
Run on a GPU:  python -m libsupermesh_amd.train_predictor --out weights.f32
The weights are LIBSUPERMESH_GPU_PREDICTOR_PARAMS float32 values in the
library's parameter order (include/libsupermesh_gpu.h).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from libsupermesh_amd import gpu, meshgen

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tet_counts_mlp.f32")


def load_weights(path=DEFAULT_WEIGHTS):
    w = np.fromfile(path, dtype=np.float32)
    if w.size != gpu.PREDICTOR_PARAMS:
        raise ValueError(f"{path}: {w.size} values, expected {gpu.PREDICTOR_PARAMS}")
    return w


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=DEFAULT_WEIGHTS)
    p.add_argument("--seeds", default="10,11")
    p.add_argument("--n", default="64,60", help="grid sizes of the two meshes")
    p.add_argument("--samples", type=int, default=5_000_000)
    p.add_argument("--epochs", type=int, default=8)
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--lr", type=float, default=2e-3)
    args = p.parse_args(argv)
    na, nb = (int(x) for x in args.n.split(","))
    seeds = [int(s) for s in args.seeds.split(",")]
    ctx = gpu.Context()
    per = args.samples // len(seeds)
    for s in seeds:
        t0 = time.time()
        pa, ea = meshgen.split_grid(3, na, s, mesh_id=0, diagonal=0)
        pb, eb = meshgen.split_grid(3, nb, s, mesh_id=1, diagonal=1)
        n = ctx.predictor_collect(gpu.DeviceMesh(pa, ea), gpu.DeviceMesh(pb, eb), per)
        print(f"[train] seed {s}: {n} samples collected ({time.time() - t0:.1f}s)", flush=True)
    t0 = time.time()
    loss = ctx.predictor_train(args.epochs, batch=args.batch, lr=args.lr, seed=1)
    print(f"[train] {args.epochs} epochs in {time.time() - t0:.1f}s, loss per epoch "
          + " ".join(f"{x:.4f}" for x in loss), flush=True)
    w = ctx.predictor_get()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    w.tofile(args.out)
    print(f"[train] wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
