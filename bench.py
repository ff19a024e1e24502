"""Objective+gradient throughput of the GPU RML imaging cost.

Workload (one GPU): 256x256 image, fine grid 512x512, ES width 7; amplitudes
on 1e5 uv points in a 10 Glambda disk, closure phases (220 triangles) and log
closure amplitudes (495 quadrangles x 2) of 12 stations x 500 timestamps;
L1 + TV + TSV + KL.  One step = one objective+gradient evaluation
(calc_cost) with x already in HBM.

Prints one JSON line.  Besides the throughput it reports:
  roofline      -- the dominant kernel's algorithmic bytes per launch over its
                   average launch duration (hipEvents on the library's stream)
  cpu_baseline  -- the FP64 CPU restatement of the same cost (C + OpenMP,
                   ES-kernel NUFFT) on a bounded sample of evaluations on
                   this host's cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def stage_bytes(kw, nf1, nf2):
    """Algorithmic HBM bytes per launch of each kernel stage."""
    import numpy as np
    Nx, Ny = kw["Nx"], kw["Ny"]
    N = np.asarray(kw["xidx"]).size
    M = np.asarray(kw["u"]).size
    G = nf1 * nf2
    nfcv = 0 if kw.get("uvidxfcv") is None else len(kw["uvidxfcv"])
    namp = 0 if kw.get("uvidxamp") is None else len(kw["uvidxamp"])
    ncp = 0 if kw.get("uvidxcp") is None else kw["uvidxcp"].shape[1]
    nca = 0 if kw.get("uvidxca") is None else kw["uvidxca"].shape[1]
    nreg = sum(1 for r in ("l1", "tv", "tsv", "kl", "gs") if kw.get(r + "_l", -1) > 0)
    half = 8 * nf1 * Ny
    return {
        "t2_rows": 8 * Nx * Ny + half,                       # x in (f64), half grid out
        "t2_cols": half + 8 * G,                             # half grid in, grid out
        "interp": 8 * G + 12 * M + 8 * M + 16 * M,           # grid, point coords, V out, Vres cleared
        "chisq": 32 * nfcv + 24 * namp + 32 * ncp + 32 * nca + 8 * M + 32 * M,  # data, V, Vres r+w
        "spread": 12 * M + 16 * M + 8 * G,                   # coords, Vres, grid out
        "t1_cols": 8 * G + 8 * G + half,                     # grid in, grid cleared, half out
        "t1_rows": half + 8 * Nx * Ny,                       # half in, grad2d out (f64)
        "reg": 8 * N * (3 + nreg) + 8 * N,                   # x, grad2d, weights in, grad out
    }


def cpu_baseline(prep, budget_s=15.0):
    """The oracle's C + OpenMP FP64 restatement of the same cost (one type-2
    and two type-1 ES-kernel NUFFTs as the reference's calc_chisq does, the
    four chi-square loops, the regularizers) timed on a bounded number of
    evaluations of the same workload on this host's cores."""
    from oracle.cpu import CpuCost
    cc = CpuCost(**prep.kwargs)
    x = prep.x0
    cc.cost_grad(x)                       # warm-up (page-in, thread pool)
    n, t0 = 0, time.perf_counter()
    while True:
        cc.cost_grad(x)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 2000:
            break
    thr = cc.threads
    cc.close()
    return {"value": round(n / el, 3), "unit": "evals/s", "cores": thr, "kind": "port",
            "sample": f"{n} full evaluations of the same workload, {el:.1f} s, C + OpenMP FP64 "
                      f"restatement (ES-kernel NUFFT, 2 type-1 transforms as in the reference) "
                      f"on {thr} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp64",
                    help="NUFFT arithmetic (chi-square and regularizers are FP64 in both)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the objective+gradient path has no multi-GPU "
              "component in this build (see DESIGN.md for how it would shard); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    torch.cuda.set_device(0)
    from smili_amd import workloads
    from smili_amd.cost import CostProblem

    t_setup = time.perf_counter()
    prep, _ = workloads.single_large(seed=args.seed)
    kw = prep.kwargs
    tune = dict(use_graph=-1 if args.no_graph else 0, precision=64 if args.precision == "fp64" else 32)
    g = CostProblem(**kw, tune=tune)
    info = g.info()
    t_setup = time.perf_counter() - t_setup
    # x goes to the problem's own device buffer once; every timed step is one
    # evaluation on resident data (the captured graph of all kernels)
    cost_host, _, _ = g.cost_grad(prep.x0)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        g.eval_async()
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.eval_async()
    g.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    value = args.steps / dt
    cost = g.cost_grad(prep.x0)[0]
    assert cost == cost_host
    info = g.info()
    # host-pointer API (x host->device, grad device->host each call): PCIe-inclusive
    nh = max(20, args.steps // 10)
    th = time.perf_counter()
    for _ in range(nh):
        g.cost_grad(prep.x0)
    host_rate = nh / (time.perf_counter() - th)

    prof = g.profile(reps=max(10, min(args.steps, 100)))
    sb = stage_bytes(kw, info["nf1"], info["nf2"])
    dom = max((k for k in prof if k in sb), key=lambda k: prof[k])
    achieved = sb[dom] / (prof[dom] * 1e-3) / 1e9
    eval_bytes = workloads.algorithmic_bytes(prep)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_" + dom, "kernel_ms": round(prof[dom], 5),
                "kernel_bytes": sb[dom],
                "eval_bytes": eval_bytes,
                "eval_achieved_GBs": round(eval_bytes / (ms * 1e-3) / 1e9, 2),
                "stage_ms": {k: round(v, 5) for k, v in prof.items() if v > 0}}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prep, args.cpu_budget)

    line = {
        "metric": "obj+grad evals/s (256^2, 1e5 vis)",
        "value": round(value, 2),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32" if args.precision == "fp32" else "f64") + " NUFFT + f64 chi-square/regularizers",
        "data": "synthetic",
        "config": {"workload": "single_large: 256x256 (fine grid 512x512, ES w=7, eps 1e-6), "
                               "1e5 amplitudes in a 10 Glambda disk + 110,000 closure phases + "
                               "495,000 log closure amplitudes (12 stations x 500 timestamps), "
                               "L1+TV+TSV+KL",
                   "npix": int(np.asarray(kw["xidx"]).size), "nuv": info["nuv"],
                   "graph": bool(info["graph"]), "setup_s": round(t_setup, 2),
                   "precision": args.precision, "cost": cost,
                   "host_api_evals_s": round(host_rate, 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
