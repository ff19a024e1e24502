"""Training of the warm-start network on the GPU:
``python -m smili_amd.warmstart.train --data DIR --out FILE.npz``.

The set (shards written by smili_amd.warmstart.gen) is uploaded once to
device memory; every step draws a batch of (sample, symmetry of the square)
pairs, so the eight symmetries multiply the set without copies.  Adam with
a linear warm-up and a cosine decay; the last ``--holdout`` fraction of the
seeds is never trained on and gives the validation error.

``--init FILE.npz`` continues from a saved network (full float32
parameters; Adam's moments start afresh), so a set too large to generate
in one session can be trained on in consecutive parts of fresh seeds, each
part generated and then trained on from the previous part's network.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import gen
from .unet import UNet


def init_params(rng):
    """He-normal 3x3 weights, zero biases, head weights N(0, 1/32)."""
    from .unet import LAYERS, nparams
    parts = []
    for _, ci, co in LAYERS:
        parts.append(rng.standard_normal(co * 9 * ci) * math.sqrt(2.0 / (9 * ci)))
        parts.append(np.zeros(co))
    parts.append(rng.standard_normal(32) * math.sqrt(1.0 / 32))
    parts.append(np.zeros(1))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == nparams()
    return p


def lr_schedule(nsteps, lr_max, warmup, lr_min_frac=0.01):
    s = np.arange(nsteps, dtype=np.float64)
    warm = np.minimum(1.0, (s + 1) / max(warmup, 1))
    cos = lr_min_frac + (1 - lr_min_frac) * 0.5 * (1 + np.cos(np.pi * np.clip((s - warmup) / max(nsteps - warmup, 1),
                                                                                  0, 1)))
    return (lr_max * warm * np.where(s < warmup, 1.0, cos)).astype(np.float32)


def validate(net, x, t):
    out = net.forward(x)
    return float(np.mean((out - t) ** 2)), float(np.linalg.norm(out - t) / max(np.linalg.norm(t), 1e-30))


def train(x, t, epochs=30, batch=32, lr=1e-3, warmup=300, holdout=0.05, seed=0, chunk=200, log=print,
          weight_decay=0.0, params=None):
    """Returns (net, report dict)."""
    n = x.shape[0]
    nval = max(1, int(round(holdout * n)))
    ntr = n - nval
    rng = np.random.default_rng(seed)
    net = UNet(batch, x.shape[1], x.shape[2], params=params if params is not None else init_params(rng))
    net.set_data(x[:ntr], t[:ntr])
    steps_per_epoch = (ntr * 8) // batch
    nsteps = epochs * steps_per_epoch
    lrs = lr_schedule(nsteps, lr, warmup)
    log(f"[train] {ntr} training samples x 8 symmetries, {nval} held out; {nsteps} steps of {batch}")
    hist, vhist = [], []
    t0 = time.time()
    step = 0
    for ep in range(epochs):
        order = rng.permutation(ntr * 8)[:steps_per_epoch * batch]
        idx = (8 * (order // 8) + (order % 8)).reshape(steps_per_epoch, batch)
        for c0 in range(0, steps_per_epoch, chunk):
            c1 = min(steps_per_epoch, c0 + chunk)
            losses = net.train(idx[c0:c1], lrs[step:step + c1 - c0], weight_decay=weight_decay)
            if not np.all(np.isfinite(losses)):
                raise FloatingPointError(f"non-finite training loss at step {step}")
            hist.extend(losses.tolist())
            step += c1 - c0
        v = validate(net, x[ntr:], t[ntr:])
        vhist.append(v)
        log(f"[train] epoch {ep + 1}/{epochs} step {step} loss {np.mean(hist[-steps_per_epoch:]):.5f} "
            f"val mse {v[0]:.5f} val rel-l2 {v[1]:.4f} lr {lrs[step - 1]:.2e} ({time.time() - t0:.0f} s)")
    rep = dict(n_train=ntr, n_val=nval, epochs=epochs, batch=batch, lr=lr, steps=nsteps, seconds=time.time() - t0,
               train_loss=np.asarray(hist, np.float32), val=np.asarray(vhist, np.float64))
    return net, rep


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", required=True, help="shard directories, comma-separated")
    ap.add_argument("--out", required=True)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--holdout", type=float, default=0.05)
    ap.add_argument("--weight-decay", type=float, default=0.0)
    ap.add_argument("--packed", action="store_true", help="write the packed (bfloat16 3x3 weights) form")
    ap.add_argument("--init", default=None, help="continue from this saved network (full float32 form)")
    a = ap.parse_args(argv)
    x, t, meta = gen.load(a.data, log=lambda s: print(s, flush=True))
    print(f"[train] loaded {x.shape[0]} samples from {a.data}", flush=True)
    params, prior = None, {}
    if a.init:
        z = np.load(a.init)
        if "params" not in z.files:
            raise SystemExit(f"{a.init} holds the packed form; continuing needs the full float32 parameters")
        params = z["params"].astype(np.float32)
        prior = {k: z[k] for k in ("seeds", "n_seen", "parts") if k in z.files}
        print(f"[train] continuing from {a.init} ({int(prior.get('n_seen', 0))} samples seen before)", flush=True)
    net, rep = train(x, t, a.epochs, a.batch, a.lr, a.warmup, a.holdout, weight_decay=a.weight_decay,
                     log=lambda s: print(s, flush=True), params=params)
    from .data import CFG
    # seeds of every part trained on so far, and how many were trained on
    # (each part's held-out tail excluded)
    seeds = np.concatenate([prior["seeds"], meta["seed"]]) if "seeds" in prior else meta["seed"]
    n_seen = int(prior.get("n_seen", 0)) + rep["n_train"]
    parts = int(prior.get("parts", 0)) + 1
    # the packed form ships; the full float32 form (beside it) continues
    outs = [(a.out, a.packed)]
    if a.packed:
        outs.append((a.out[:-4] + "_full.npz", False))
    for out, packed in outs:
        net.save(out, packed=packed, cfg=json.dumps(CFG), seeds=seeds, n_train=rep["n_train"], n_seen=n_seen,
                 parts=parts, epochs=rep["epochs"], train_loss=rep["train_loss"][::max(1, rep["steps"] // 2000)],
                 val=rep["val"], seconds=rep["seconds"])
        print(f"[train] saved {out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
