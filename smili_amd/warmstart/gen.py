"""Training-set generation: ``python -m smili_amd.warmstart.gen --seeds
A:B --workers K --out DIR [--niter N] [--inputs] [--deadline S]``.

The driver starts K worker processes (each a fresh interpreter with one
BLAS / OpenMP thread, since K workers share the host's cores; the driver
itself never touches the GPU) and reports their progress; worker i images
the seeds A + i, A + i + K, ... with the plain GPU imager
(smili_amd.warmstart.data.sample) and writes DIR/shard_i.npz (targets as
float16 and per-sample metadata, compressed), rewritten every 50 samples,
and one progress line per sample to DIR/shard_i.log.  Network inputs are
a deterministic function of the seed (dirty maps, milliseconds on the GPU)
and are recomputed by ``load`` unless ``--inputs`` stores them (float16)
in the shards.  ``--niter`` caps the imager's L-BFGS-B iterations (a
target then stops short of convergence: from the flat prior the objective
is within ~1e-3 of its final value after ~90 iterations, far inside the
network's own error); ``--deadline`` makes the workers stop taking seeds
that many seconds after they start (a set bounded by time, not by count).
"""
from __future__ import annotations

import argparse
import glob
import os
import subprocess
import sys
import time

import numpy as np

KINDS = ("ring", "crescent", "gaussians", "disk")
META = ("seed", "nvis", "niter", "nfev", "cost", "solve_s", "total_s", "flux", "dirty_peak", "target_peak")


def _save(path, ts, metas, xs=None):
    out = {"t": np.asarray(ts, np.float16), "kind": np.array([KINDS.index(m["kind"]) for m in metas], np.int8)}
    if xs is not None:
        out["x"] = np.asarray(xs, np.float16)
    for k in META:
        out[k] = np.array([m[k] for m in metas])
    tmp = path + ".tmp.npz"
    np.savez_compressed(tmp, **out)
    os.replace(tmp, path)


def worker(seeds, shard, niter=None, inputs=False, deadline=None):
    t0 = time.time()
    log = open(shard[:-4] + ".log", "a", buffering=1)
    from .data import CFG, sample
    cfg = dict(CFG) if niter is None else dict(CFG, niter=int(niter))
    log.write(f"# started, imports {time.time() - t0:.1f} s\n")
    ts, metas, xs = [], [], ([] if inputs else None)
    for i, s in enumerate(seeds):
        if deadline is not None and time.time() - t0 > deadline:
            log.write(f"# deadline after {i} samples\n")
            break
        x, t, m = sample(int(s), cfg, log=(lambda msg: log.write(msg + "\n")) if i < 2 else None)
        ts.append(t)
        metas.append(m)
        if inputs:
            xs.append(x)
        log.write(f"{s} {m['kind']} nvis={m['nvis']} niter={m['niter']} solve={m['solve_s']:.2f}s\n")
        if (i + 1) % 50 == 0:
            _save(shard, ts, metas, xs)
    _save(shard, ts, metas, xs)
    log.close()


def load(dirnames, inputs=True, log=None):
    """All shards of one or more directories, ordered by seed: (x [N, n, n,
    2] float32 recomputed from the seeds -- None unless ``inputs`` --,
    t [N, n, n] float32, metadata dict of arrays)."""
    if isinstance(dirnames, str):
        dirnames = [d for d in dirnames.split(",") if d]
    paths = sorted(p for d in dirnames for p in glob.glob(os.path.join(d, "shard_*.npz")))
    if not paths:
        raise FileNotFoundError(f"no shard_*.npz in {dirnames}")
    parts = [np.load(p) for p in paths]
    seed = np.concatenate([p["seed"] for p in parts])
    o = np.argsort(seed, kind="stable")
    t = np.concatenate([p["t"] for p in parts])[o].astype(np.float32)
    meta = {k: np.concatenate([p[k] for p in parts])[o] for k in ("kind",) + META}
    if len(np.unique(meta["seed"])) != len(meta["seed"]):
        raise ValueError("duplicate seeds across shards")
    x = None
    if inputs and all("x" in p.files for p in parts):
        x = np.concatenate([p["x"] for p in parts])[o].astype(np.float32)
    elif inputs:
        from .data import dirty_maps, make_case
        x = np.empty(t.shape + (2,), np.float32)
        t0 = time.time()
        for i, s in enumerate(meta["seed"]):
            d, b, _ = dirty_maps(make_case(int(s)))
            x[i, ..., 0], x[i, ..., 1] = d, b
            if log and (i + 1) % 500 == 0:
                log(f"[load] inputs {i + 1}/{len(t)} ({time.time() - t0:.0f} s)")
    return x, t, meta


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", required=True, help="A:B")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", required=True)
    ap.add_argument("--niter", type=int, default=None, help="cap of the imager's iterations per target")
    ap.add_argument("--inputs", action="store_true", help="store the network inputs in the shards")
    ap.add_argument("--deadline", type=float, default=None, help="seconds after which workers stop")
    ap.add_argument("--worker", type=int, default=-1, help=argparse.SUPPRESS)
    a = ap.parse_args(argv)
    lo, hi = (int(s) for s in a.seeds.split(":"))
    if a.worker >= 0:
        worker(range(lo + a.worker, hi, a.workers), os.path.join(a.out, f"shard_{a.worker:02d}.npz"),
               niter=a.niter, inputs=a.inputs, deadline=a.deadline)
        return 0
    if not 1 <= a.workers <= 16:
        raise SystemExit("--workers must be in 1..16 (processes that share the GPU)")
    os.makedirs(a.out, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    extra = (["--niter", str(a.niter)] if a.niter is not None else []) + (["--inputs"] if a.inputs else []) + \
        (["--deadline", str(a.deadline)] if a.deadline is not None else [])
    procs = [subprocess.Popen([sys.executable, "-m", "smili_amd.warmstart.gen", "--seeds", a.seeds,
                               "--workers", str(a.workers), "--out", a.out, "--worker", str(i)] + extra,
                              stdout=open(os.path.join(a.out, f"worker_{i:02d}.out"), "w"),
                              stderr=subprocess.STDOUT, env=env)
             for i in range(a.workers)]
    t0 = time.time()
    while any(p.poll() is None for p in procs):
        time.sleep(15)
        done = 0
        for i in range(a.workers):
            try:
                with open(os.path.join(a.out, f"shard_{i:02d}.log")) as f:
                    done += sum(1 for ln in f if not ln.startswith("#"))
            except FileNotFoundError:
                pass
        print(f"[gen] {done}/{hi - lo} samples after {time.time() - t0:.0f} s", flush=True)
        if any(p.poll() not in (None, 0) for p in procs):
            for p in procs:
                if p.poll() is None:
                    p.terminate()
            break
    codes = [p.wait() for p in procs]
    print(f"[gen] workers exited {codes} after {time.time() - t0:.0f} s", flush=True)
    return 0 if all(c == 0 for c in codes) else 1


if __name__ == "__main__":
    sys.exit(main())
