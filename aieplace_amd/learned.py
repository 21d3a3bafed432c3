"""Learned Nesterov step length: record collection, GPU training, weights I/O.

The model (16 -> 64 -> 64 -> 1, GELU; include/eplace_mlp.h) is trained by the
library's own HIP kernels (ep_mlp_train) on records (features, log10(accepted
step / bin width)) that ep_step collects during plain placements.  In a
learned run each proposal is verified by the backtracking condition and falls
back to the Lipschitz search when it fails.

Usage (on a GPU):  python -m aieplace_amd.learned --runs 200 --out <file.npy>
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

from . import api
from .configs import CONFIGS, bins_for_cells, placer_params
from .netlist import generate
from .placer import Placer

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "step_mlp_weights.npy")


def training_instances(runs: int, lo: int = 50_000, hi: int = 500_000, seed: int = 1000):
    """(n_cells, seed) pairs, cell counts log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi), runs)).astype(int)
    return [(int(n), seed + 1 + k) for k, n in enumerate(ns)]


def collect_records(instances, precision=api.EP_PREC_BF16X3, log=print):
    feats, targets = [], []
    cfg = CONFIGS["stdcell_2m"]
    for k, (n, seed) in enumerate(instances):
        nl = generate("stdcell", seed=seed, n_movable=n)
        # the grid that the workloads pair with this size: 0.3-1.5 cells per bin
        b = bins_for_cells(n) if n < 700_000 else 1024
        p = placer_params(dict(cfg, bins=b), precision=precision, use_graph=1)
        with Placer(nl, p) as pl:
            res = pl.run()
            f, t = pl.take_step_records()
        feats.append(f)
        targets.append(t)
        log(f"  run {k}: {n} cells, bins {b}, {res['iterations']} it, {len(t)} records, "
            f"{res['seconds']:.3f}s")
    return np.concatenate(feats), np.concatenate(targets)


def train(features: np.ndarray, log_step: np.ndarray, epochs: int = 40, batch: int = 512,
          lr: float = 2e-3, seed: int = 1):
    tp = api.EpMlpTrainParams()
    api.lib.ep_mlp_default_train_params(C.byref(tp))
    tp.epochs, tp.batch, tp.lr, tp.seed = epochs, batch, lr, seed
    X = np.ascontiguousarray(features, np.float32)
    Y = np.ascontiguousarray(log_step, np.float32)
    w = np.zeros(api.MLP_NWEIGHTS, np.float32)
    curve = np.zeros(epochs, np.float32)
    api.check(api.lib.ep_mlp_train(api.fptr(X), api.fptr(Y), len(Y), C.byref(tp), api.fptr(w),
                                   api.fptr(curve)))
    return w, curve


def forward(weights, features):
    X = np.ascontiguousarray(features, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    out = np.zeros(len(X), np.float32)
    api.check(api.lib.ep_mlp_forward(api.fptr(w), api.fptr(X), len(X), api.fptr(out)))
    return out


def grad(weights, features, log_step):
    X = np.ascontiguousarray(features, np.float32)
    Y = np.ascontiguousarray(log_step, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    g = np.zeros(api.MLP_NWEIGHTS, np.float32)
    loss = C.c_float()
    api.check(api.lib.ep_mlp_grad(api.fptr(w), api.fptr(X), api.fptr(Y), len(Y), api.fptr(g),
                                  C.byref(loss)))
    return g, loss.value


def save(weights, path=DEFAULT_WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(weights, np.float32), allow_pickle=False)


def load(path=DEFAULT_WEIGHTS):
    w = np.load(path, allow_pickle=False).astype(np.float32)
    if w.shape != (api.MLP_NWEIGHTS,):
        raise ValueError(f"{path}: expected {api.MLP_NWEIGHTS} weights, got {w.shape}")
    return w


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=200)
    ap.add_argument("--lo", type=int, default=50_000)
    ap.add_argument("--hi", type=int, default=500_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args()
    t0 = time.time()
    X, Y = collect_records(training_instances(args.runs, args.lo, args.hi),
                           log=lambda s: print(s, flush=True))
    print(f"{len(Y)} records from {args.runs} runs in {time.time() - t0:.1f}s", flush=True)
    t0 = time.time()
    w, curve = train(X, Y, epochs=args.epochs)
    print(f"trained in {time.time() - t0:.1f}s; loss per epoch: "
          + " ".join(f"{v:.4f}" for v in curve), flush=True)
    save(w, args.out)
    print(f"saved {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
