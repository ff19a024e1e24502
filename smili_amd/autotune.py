"""Autotuner of the cost kernels' work sizes (``smili_tune``).

``autotune(make, x0)`` times one objective+gradient evaluation of the handle
``make(tune)`` builds, for candidate ``smili_tune`` settings, and returns the
fastest setting with every candidate's time.  The search is coordinate
descent: the parameters in ``SPACE`` order, each tried at its candidate
values on top of the best setting so far.  Every candidate must reproduce
the default setting's cost (relative ``rtol``): the settings change how the
work is split, never what is computed, so a different cost would be a bug
rather than a speed-up.  Settings the library rejects (an LDS budget
exceeded, say) are recorded and skipped.

The defaults in the library are what this search finds on the bench
workloads on MI355X; the tuner is for other problem shapes:
``imaging(..., tune="auto")`` runs it before a solve, and
``python -m smili_amd.autotune [--workload single_large|movie|survey]``
prints the search for a bench workload.
"""
from __future__ import annotations

import time

import numpy as np

# (smili_tune field, candidate values; the first is the library default)
SPACE = (
    ("chunk", (4096, 2048, 8192)),
    ("bin", (16, 8, 32)),
    ("sub_points", (256, 128, 512)),
    ("col_group", (0, 1, 2, 4)),
    ("fft_radix", (4, 8)),
    ("chisq_threads", (256, 512)),
    ("dft_shapes", (0, 1, 2, 3)),
)


def time_eval(g, x0, reps=100, warmup=5):
    """Mean seconds of one evaluation on the handle's own buffers (x0
    uploaded once; launches queued back to back, one synchronisation)."""
    g.cost_grad(x0)
    for _ in range(warmup):
        g.eval_async()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.eval_async()
    g.synchronize()
    return (time.perf_counter() - t0) / reps


def _cost(c):
    return float(np.sum(c))


def autotune(make, x0, space=SPACE, reps=100, rtol=None, timer=None, log=None, base=None):
    """Returns (best tune dict, rows); rows = one dict per candidate with
    "tune", "seconds" (None if rejected) and "error" (the rejection).

    make(tune) -> handle (CostProblem, MovieProblem or BatchProblem);
    timer(handle, x0) -> seconds per evaluation (default ``time_eval``);
    base: settings every candidate keeps (e.g. {"precision": 32});
    rtol: allowed relative cost difference (default 1e-9, 1e-6 in FP32)."""
    from ._lib import SmiliError
    timer = timer or (lambda g, x: time_eval(g, x, reps))
    base = dict(base or {})
    if rtol is None:
        rtol = 1e-6 if base.get("precision") == 32 else 1e-9
    rows, seen = [], {}

    def trial(tune):
        key = tuple(sorted(tune.items()))
        if key in seen:
            return seen[key]
        row = {"tune": dict(tune), "seconds": None, "error": None}
        try:
            g = make(dict(tune))
        except SmiliError as e:
            row["error"] = str(e)
        else:
            try:
                c = _cost(g.cost_grad(x0)[0])
                if ref_cost[0] is not None and abs(c - ref_cost[0]) > rtol * abs(ref_cost[0]):
                    raise RuntimeError(f"tune {tune} changes the cost: {c!r} vs {ref_cost[0]!r}")
                if ref_cost[0] is None:
                    ref_cost[0] = c
                row["seconds"] = float(timer(g, x0))
            finally:
                g.close()
        rows.append(row)
        seen[key] = row
        if log:
            log(row)
        return row

    ref_cost = [None]
    first = trial(base)
    if first["seconds"] is None:
        raise RuntimeError(f"the default setting is rejected: {first['error']}")
    best, best_s = dict(base), first["seconds"]
    for name, values in space:
        for v in values:
            cand = dict(best)
            cand[name] = v
            r = trial(cand)
            if r["seconds"] is not None and r["seconds"] < best_s:
                best, best_s = cand, r["seconds"]
    return best, rows


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="autotune the cost kernels on a bench workload")
    ap.add_argument("--workload", default="single_large", choices=("single_large", "movie", "survey"))
    ap.add_argument("--reps", type=int, default=100)
    a = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    from . import workloads
    from .cost import BatchProblem, CostProblem, MovieProblem
    if a.workload == "single_large":
        prep, _ = workloads.single_large()
        make, x0 = (lambda t: CostProblem(**prep.kwargs, tune=t)), prep.x0
    elif a.workload == "movie":
        prep, _ = workloads.dynamic_movie()
        make, x0 = (lambda t: MovieProblem(**prep.kwargs, tune=t)), prep.x0
    else:
        preps, _ = workloads.survey()
        kws = [p.kwargs for p in preps]
        make, x0 = (lambda t: BatchProblem(kws, tune=t)), np.concatenate([p.x0 for p in preps])

    def log(r):
        ms = None if r["seconds"] is None else round(r["seconds"] * 1e3, 4)
        print(json.dumps({"tune": r["tune"], "ms": ms, "error": r["error"]}), flush=True)

    best, rows = autotune(make, x0, reps=a.reps, log=log)
    print(json.dumps({"workload": a.workload, "best": best,
                      "ms": round(min(r["seconds"] for r in rows if r["seconds"]) * 1e3, 4),
                      "default_ms": round(rows[0]["seconds"] * 1e3, 4)}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
