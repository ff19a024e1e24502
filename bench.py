"""Objective+gradient throughput of the GPU RML imaging cost.

Default workload (one GPU): 256x256 image, fine grid 512x512, ES width 7;
amplitudes on 1e5 uv points in a 10 Glambda disk, closure phases (220
triangles) and log closure amplitudes (495 quadrangles x 2) of 12 stations x
500 timestamps; L1 + TV + TSV + KL.  One step = one objective+gradient
evaluation (calc_cost) with x already in HBM.

``--workload movie``: 64 frames of 128x128, 3e5 visibilities, per-frame
chi-square + L1 + TSV + Rt as one joint cost (fftim3d.calc_cost); one step =
one evaluation of the whole movie.  ``--workload survey``: 512 independent
problems of 64^2..256^2 evaluated together; one step = one evaluation of all
512 (the value counts problem evaluations per second).  ``--workload
warmstart``: time-to-image of the learned warm start against the plain
L-BFGS-B imager on held-out 128^2 cases (L1 + TSV); one step = one case,
the value is the median speed-up, and the roofline is the MFMA one of the
network's dominant training kernel.  ``--workload warmstart_movie``: the same
for the 64-frame movie with the network's proposal applied per frame; one
step = one movie (plain and warm solves), the roofline is the movie cost's
dominant kernel.

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
FP64_PEAK_TFLOPS = 78.6        # MI355X dense FP64 (matrix and vector)


def fine_size(n):
    """Fine-grid length of an n-pixel axis: the smallest 2,3,5-smooth
    integer >= max(2 n, 14) (upsampling 2, ES width 7)."""
    m = max(2 * n, 14)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def stage_bytes(kws, units, fp64=True):
    """Algorithmic bytes per kernel stage of one evaluation: every input the
    stage's algorithm needs read once and every output written once, in the
    precision it computes in (complex grid, model visibility and DFT-matrix
    elements of 16 bytes in FP64, 8 in FP32; ES weights 14 x 8 or 14 x 4 bytes
    per point), with the data records of the cost model (index + datum +
    variance: 28 B per full visibility or closure phase, 20 per amplitude,
    32 per log closure amplitude -- or the stored record where it is smaller:
    24 B per log closure amplitude, whose four point indices are 16-bit
    chunk-local slots) and one FP64 complex residual per point (plus the
    spreading's 16-byte per-point reference record where a point's
    residual shares must be gathered, i.e. unless every point has one
    datum); a regularizer's per-pixel weights count only where they vary over the
    image (uniform weights -- the flat-prior TV / TSV -- are one value the
    kernels do not read per pixel).
    ``kws``: the problems' cost keyword dicts (a movie's dict carries one
    entry per frame in ``units``); ``units``: the handle's unit_info() of
    each unit in the same order (transform, grid or window width / height)."""
    import numpy as np
    cb = 16 if fp64 else 8                 # complex grid element / model visibility
    wb = 2 * 7 * (8 if fp64 else 4)        # ES weights per point
    tot = {}

    def add(k, v):
        tot[k] = tot.get(k, 0) + int(v)

    ui = 0
    def n_of(kw, key, k=1):
        a = kw.get(key)
        return 0 if a is None else np.asarray(a).size // k

    def one_datum_per_point(kw):
        idx = kw.get("uvidxfcv") if n_of(kw, "uvidxfcv") else kw.get("uvidxamp")
        if idx is None or n_of(kw, "uvidxcp", 3) or n_of(kw, "uvidxca", 4) or \
                (n_of(kw, "uvidxfcv") and n_of(kw, "uvidxamp")):
            return False
        a = np.abs(np.asarray(idx)).ravel()
        return a.size == np.asarray(kw["u"]).size and np.unique(a).size == a.size
    # the plan reads each point's residual share at the point's own index,
    # with no per-point reference record, when every point has one datum
    pref_bytes = 0 if all(one_datum_per_point(kw) for kw in kws) else 16
    for kw in kws:
        N = np.asarray(kw["xidx"]).size

        def varies(w):
            w = np.asarray(w)
            return w.size > 1 and not np.all(w == w.flat[0])
        # per-pixel weight arrays the regularizer launch reads
        nreg = sum(1 for r in ("l1", "tv", "tsv", "kl", "gs", "rt", "ri")
                   if kw.get(r + "_l", -1) > 0 and varies(kw.get(r + "_wgt")))
        M_all = 0
        for u in kw.get("units", [kw]):
            info = units[ui]
            ui += 1
            Nx, Ny, M = info["nx"], info["ny"], info["m"]
            M_all += M
            Npx = N                                    # pixels per image (frame)
            if info["transform"] == 1:
                G = info["nf1"] * info["nf2"]
                half = cb * info["nf1"] * Ny
                add("t2_rows", 8 * Nx * Ny + half)
                add("t2_cols", half + cb * G)
                add("t1_cols", 2 * cb * G + half)
                add("t1_rows", half + 16 * Npx)
            elif info["transform"] == 2:
                gw, gh = info["gw"], info["gh"]
                G = gw * gh
                add("dft2_y", 8 * Nx * Ny + cb * gh * Ny + cb * gh * Nx)
                add("dft2_x", cb * gh * Nx + cb * Nx * gw + cb * G)
                add("dft1_x", cb * G + cb * gw * Nx + cb * gh * Nx)
                add("dft1_y", cb * Ny * gh + cb * gh * Nx + 16 * Npx + cb * G)
            else:
                G = 0
            if M:
                add("interp", cb * G + M * (4 + wb + cb))
                add("spread", M * (4 + wb + pref_bytes + 16) + cb * G)   # cell, weights, slot refs, residual; grid
        nun = len(kw.get("units", [kw]))
        add("reg", 8 * N * nun * (1 + nreg) + 8 * N * nun)         # x, weights in, gradreg out
        nfcv = 0 if kw.get("uvidxfcv") is None else len(kw["uvidxfcv"])
        namp = 0 if kw.get("uvidxamp") is None else len(kw["uvidxamp"])
        ncp = 0 if kw.get("uvidxcp") is None else kw["uvidxcp"].shape[1]
        nca = 0 if kw.get("uvidxca") is None else kw["uvidxca"].shape[1]
        add("chisq", 28 * nfcv + 20 * namp + 28 * ncp + 24 * nca + cb * M_all + 16 * M_all)
    return tot


def stage_flops(units):
    """FP64 FLOP of the pruned-DFT products per evaluation (complex
    multiply-add = 8 FLOP, complex x real = 4): S = Ey I (I real), G = S Ex,
    P = Gs ExT, Re(EyT P) (two real products).  Units on the FFT path add
    nothing here."""
    tot = {}
    for info in units:
        if info["transform"] != 2:
            continue
        nx, ny, gw, gh = info["nx"], info["ny"], info["gw"], info["gh"]
        for k, v in (("dft2_y", 4 * gh * ny * nx), ("dft2_x", 8 * gh * nx * gw),
                     ("dft1_x", 8 * gh * gw * nx), ("dft1_y", 4 * ny * gh * nx)):
            tot[k] = tot.get(k, 0) + v
    return tot


def survey_bytes(kws):
    """The cost model's bytes per evaluation (complex64 grids, two passes
    per transform, FP32 uv pairs read by both transforms, complex64 model
    visibilities written and read, FP64 residuals written and read, the
    data records read once): the figure the HBM ceiling of the headline
    metric is quoted on."""
    import numpy as np
    b = 0
    for kw in kws:
        N = np.asarray(kw["xidx"]).size
        for u in kw.get("units", [kw]):
            M = u["M"] if "M" in u else np.asarray(kw["u"]).size
            G = fine_size(u["Nx"]) * fine_size(u["Ny"])
            b += 32 * G + 16 * M + 16 * M + 32 * M
        b += 16 * N
        nfcv = 0 if kw.get("uvidxfcv") is None else len(kw["uvidxfcv"])
        namp = 0 if kw.get("uvidxamp") is None else len(kw["uvidxamp"])
        ncp = 0 if kw.get("uvidxcp") is None else kw["uvidxcp"].shape[1]
        nca = 0 if kw.get("uvidxca") is None else kw["uvidxca"].shape[1]
        b += 28 * nfcv + 20 * namp + 28 * ncp + 32 * nca
    return int(b)


def _movie_units(kw):
    return [dict(Nx=kw["Nx"], Ny=kw["Ny"], M=int(m)) for m in kw["Nuvs"]]


def _timed_loop(g, steps, warmup):
    """Average wall time of `steps` evaluations on resident buffers."""
    import torch
    for _ in range(warmup):
        g.eval_async()
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.eval_async()
    g.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _cpu_time_per_eval(kw, budget_s, max_evals=2000):
    """Seconds per evaluation of the oracle's C + OpenMP FP64 port of the
    2-D cost on one problem, and the thread count."""
    from oracle.cpu import CpuCost
    import numpy as np
    cc = CpuCost(**{k: v for k, v in kw.items() if k not in ("x0",)})
    x = kw["x0"]
    cc.cost_grad(x)                       # warm-up (page-in, thread pool)
    n, t0 = 0, time.perf_counter()
    while True:
        cc.cost_grad(x)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= max_evals:
            break
    thr = cc.threads
    cc.close()
    return el / n, n, el, thr


def cpu_baseline(prep, budget_s=15.0):
    """The oracle's C + OpenMP FP64 restatement of the same cost (one type-2
    and two type-1 ES-kernel NUFFTs as the reference's calc_chisq does, the
    four chi-square loops, the regularizers) timed on a bounded number of
    evaluations of the same workload on this host's cores."""
    per, n, el, thr = _cpu_time_per_eval(dict(prep.kwargs, x0=prep.x0), budget_s)
    return {"value": round(1.0 / per, 3), "unit": "evals/s", "cores": thr, "kind": "port",
            "sample": f"{n} full evaluations of the same workload, {el:.1f} s, C + OpenMP FP64 "
                      f"restatement (ES-kernel NUFFT, 2 type-1 transforms as in the reference) "
                      f"on {thr} threads"}


def _frame_kwargs(kw, x0, z):
    """Frame z of a movie as a 2-D cost (its uv slice and the full complex
    visibilities on it, re-indexed; static regularizers at 1/Nz)."""
    import numpy as np
    nuvs = np.asarray(kw["Nuvs"])
    s = int(nuvs[:z].sum())
    e = s + int(nuvs[z])
    npix = np.asarray(kw["xidx"]).size
    idx = np.asarray(kw["uvidxfcv"])
    sel = (np.abs(idx) > s) & (np.abs(idx) <= e)
    f = {k: kw[k] for k in ("Nx", "Ny", "xidx", "yidx", "Nxref", "Nyref")}
    f.update(u=kw["u"][s:e], v=kw["v"][s:e], uvidxfcv=np.sign(idx[sel]) * (np.abs(idx[sel]) - s),
             Vfcv=kw["Vfcv"][sel], Varfcv=kw["Varfcv"][sel], wfcv=kw["wfcv"])
    for r in ("l1", "tsv", "tv", "kl", "gs"):
        if kw.get(r + "_l", -1) > 0:
            f[r + "_l"] = kw[r + "_l"] / len(nuvs)
            f[r + "_wgt"] = kw[r + "_wgt"]
    f["x0"] = x0[z * npix:(z + 1) * npix]
    return f


def cpu_baseline_movie(prep, budget_s=15.0, n_sample=8):
    """CPU port on a sample of frames: frame z's chi-square and static
    regularizers are the 2-D cost of its own data (the movie's data here are
    per-frame visibilities; Rt adds one pass over the frames), so one movie
    evaluation costs the sum over frames; n_sample evenly spaced frames are
    timed and the mean scaled to all frames."""
    import numpy as np
    kw = prep.kwargs
    Nz = len(kw["Nuvs"])
    zs = np.linspace(0, Nz - 1, n_sample).round().astype(int)
    per = []
    thr = 1
    for z in zs:
        t, _, _, thr = _cpu_time_per_eval(_frame_kwargs(kw, prep.x0, z), budget_s / n_sample)
        per.append(t)
    tot = float(np.mean(per)) * Nz
    return {"value": round(1.0 / tot, 4), "unit": "evals/s", "cores": thr, "kind": "port",
            "sample": f"C + OpenMP FP64 2-D cost of {n_sample} of the {Nz} frames (each frame's "
                      f"visibilities, L1 + TSV), {budget_s:.0f} s in all, mean per frame "
                      f"{np.mean(per) * 1e3:.2f} ms scaled by {Nz} frames; Rt (one pass) not timed"}


def cpu_baseline_survey(preps, budget_s=15.0, n_sample=16, seed=0):
    """CPU port on a random sample of the survey's problems, scaled to all."""
    import numpy as np
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(preps), n_sample, replace=False)
    per = []
    thr = 1
    for i in pick:
        t, _, _, thr = _cpu_time_per_eval(dict(preps[i].kwargs, x0=preps[i].x0), budget_s / n_sample)
        per.append(t)
    tot = float(np.mean(per)) * len(preps)
    return {"value": round(len(preps) / tot, 3), "unit": "problem-evals/s", "cores": thr, "kind": "port",
            "sample": f"C + OpenMP FP64 cost of {n_sample} random problems of the {len(preps)}, "
                      f"{budget_s:.0f} s in all, mean {np.mean(per) * 1e3:.2f} ms per problem "
                      f"scaled to the whole survey"}


COPY_CAL_BYTES = 512 * 2 ** 20        # calibration copy: 512 MiB read, 512 MiB written


def traffic_file(workload, precision):
    return os.path.join(ROOT, "profiles", f"traffic_{workload}_{precision}.json")


def _pmc_rows(path):
    """(kernel name, counter value) per dispatch from a rocprofv3
    counter_collection.csv."""
    import csv
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((r["Kernel_Name"], r.get("Counter_Name", ""), float(r["Counter_Value"])))
    return out


def ingest_pmc(fetch_csv, write_csv, out_json, workload, precision):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of the library's kernels,
    each scaled by what the same counter reads for the calibration copy of
    COPY_CAL_BYTES (an elementwise x * 1 over FP64, one read and one write
    per byte), -> HBM bytes per launch of every kernel."""
    import numpy as np
    res = {}
    for label, path in (("fetch", fetch_csv), ("write", write_csv)):
        per = {}
        for name, _, val in _pmc_rows(path):
            per.setdefault(name, []).append(val)
        res[label] = per
    cal = {}
    for label in ("fetch", "write"):
        cands = {k: v for k, v in res[label].items() if "smili" not in k and "Mul" in k}
        if not cands:
            raise SystemExit(f"no calibration copy kernel in the {label} pass")
        name = max(cands, key=lambda k: np.median(cands[k]))
        cal[label] = {"kernel": name[:120], "kib": float(np.median(cands[name])),
                      "factor": COPY_CAL_BYTES / 1024.0 / float(np.median(cands[name]))}
    kernels = {}
    for name in res["fetch"]:
        if "smili" not in name:
            continue
        short = name.split("(")[0].replace("void ", "").replace("smili::", "").split("<")[0].strip()
        f = float(np.mean(res["fetch"][name])) * cal["fetch"]["factor"] * 1024
        w = float(np.mean(res["write"].get(name, [0.0]))) * cal["write"]["factor"] * 1024
        k = kernels.setdefault(short, {"fetch_bytes": 0.0, "write_bytes": 0.0, "dispatches": 0})
        k["fetch_bytes"] += f
        k["write_bytes"] += w
        k["dispatches"] = len(res["fetch"][name])
    for k in kernels.values():
        k["traffic_bytes"] = int(k["fetch_bytes"] + k["write_bytes"])
        k["fetch_bytes"] = int(k["fetch_bytes"])
        k["write_bytes"] = int(k["write_bytes"])
    doc = {"workload": workload, "precision": precision, "calibration": cal, "kernels": kernels}
    with open(out_json, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc, indent=1))


def _traffic(workload, precision, kernel):
    path = traffic_file(workload, precision)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        doc = json.load(f)
    k = doc.get("kernels", {}).get(kernel)
    return None if k is None else k["traffic_bytes"]


def _roofline(g, kws, fp64, ms, steps, workload="single_large"):
    """The dominant kernel's algorithmic bytes per launch over its launch
    duration: hipEvents on the library's stream around 100 back-to-back
    launches of each stage (one pass), so the events' own overhead does not
    count against the kernel."""
    nunits = g.nprob if getattr(g, "nprob", 1) > 1 else getattr(g, "Nz", 1)
    units = [g.unit_info(u) for u in range(nunits)]
    prof = g.profile(reps=100, back_to_back=True)
    sb = stage_bytes(kws, units, fp64)
    eval_bytes = int(sum(sb.values()))
    names = {k: "k_" + k for k in sb}
    for k in ("dft2_y", "dft2_x", "dft1_x", "dft1_y"):
        names[k] = "k_dft"
    if prof.get("reg", 0.0) == 0.0 and prof.get("chisq", 0.0) > 0.0:
        # one-stream schedule: the regularizers ride on the chi-square launch
        # and the final sums on the spreading launch
        sb["chisq"] += sb.pop("reg")
        names["chisq"], names["spread"] = "k_chisq_reg", "k_spread_fin"
    dom = max((k for k in prof if k in sb), key=lambda k: prof[k])
    achieved = sb[dom] / (prof[dom] * 1e-3) / 1e9
    sv = survey_bytes(kws)
    sf = stage_flops(units)
    if dom in sf and fp64:
        # the DFT products are FP64 FMA chains: their roofline is the FP64
        # rate (their algorithmic bytes over the same time stay beside it)
        tf = sf[dom] / (prof[dom] * 1e-3) / 1e12
        head = {"bound": "mfma", "achieved": round(tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tf / FP64_PEAK_TFLOPS, 5), "kernel_flop": sf[dom],
                "hbm_achieved_GBs": round(achieved, 2), "hbm_frac": round(achieved / HBM_PEAK_GBS, 5)}
    else:
        head = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5)}
    return {**head,
            "traffic": _traffic(workload, "fp64" if fp64 else "fp32", names[dom]),
            "kernel": names[dom], "stage": dom, "kernel_ms": round(prof[dom], 5),
            "kernel_bytes": sb[dom],
            "eval_bytes": eval_bytes,
            "eval_achieved_GBs": round(eval_bytes / (ms * 1e-3) / 1e9, 2),
            "model_bytes": sv, "model_achieved_GBs": round(sv / (ms * 1e-3) / 1e9, 2),
            "model_ceiling_evals_s": round(HBM_PEAK_GBS * 1e9 / sv, 1),
            "stage_ms": {k: round(v, 5) for k, v in prof.items() if v > 0},
            "stage_bytes": sb, "stage_flop": sf}


MFMA_BF16_PEAK_TFLOPS = 2500.0


def _unet_roofline(params, batch=32, reps=20):
    """MFMA roofline of the network's dominant kernel launch in a training
    step (batch 32, 128^2): every 3x3 layer's forward, input-gradient and
    weight-gradient launch timed with HIP events on the handle's stream."""
    import numpy as np
    from smili_amd.warmstart.unet import LAYERS, UNet
    net = UNet(batch, 128, 128, params=params)
    x = np.random.default_rng(0).standard_normal((batch, 128, 128, 2)).astype(np.float32)
    net.set_data(x, np.zeros((batch, 128, 128), np.float32))
    step_ms, fwd_ms = net.time_steps(10)
    step_flop, fwd_flop = net.flops()
    level = (0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0)
    # kernel instantiations as the library picks them (smili_unet.hip
    # launch_conv, wgrad_plan): 8 x 16 forward / input-gradient tiles with 64
    # output channels per work-group when that gives at least 1536
    # work-groups, else 32; 8 x 32 weight-gradient tiles where the width allows
    launches, kernels = [], {}
    for li in range(1, len(LAYERS)):
        name, cin, cout = LAYERS[li]
        w = 128 >> level[li]
        tiles = batch * (w // 8) * (w // 16)
        hw = batch * w * w
        # activation bytes each launch must move at least (bf16 activations,
        # fp32 input gradients); decoder inputs [up(x1) ++ x2] read x1 at
        # half resolution
        c2 = cout if li in (8, 10, 12) else cin
        x_bytes = 2 * (hw * c2 + hw // 4 * (cin - c2))
        dp_bytes = 2 * hw * cout
        min_bytes = {0: x_bytes + 2 * hw * cout, 1: dp_bytes + 4 * hw * cin, 2: dp_bytes + x_bytes}
        for ps, pname in ((0, "forward"), (1, "input_grad"), (2, "weight_grad")):
            ms, fl = net.time_layer(li, ps, reps)
            # the time a launch cannot beat: its FLOP at the BF16 MFMA peak or
            # its bytes at the HBM peak, whichever is longer
            bound_ms = max(fl / (MFMA_BF16_PEAK_TFLOPS * 1e9), min_bytes[ps] / (HBM_PEAK_GBS * 1e6))
            launches.append((ms, fl, name, pname, bound_ms))
            # the kernel instantiation this launch runs (the library's choice)
            if ps == 2:
                kern = f"k_conv3x3_wgrad<{4 if cout % 64 == 0 else 2}, {32 if w % 32 == 0 else 16}>"
            else:
                co = cout if ps == 0 else cin
                kern = f"k_conv3x3<16, {4 if co % 64 == 0 and tiles * (co // 64) >= 1536 else 2}>"
            k = kernels.setdefault(kern, [0.0, 0.0, 0, 0.0])
            k[0] += ms
            k[1] += fl
            k[2] += 1
            k[3] += bound_ms
    net.close()
    # dominant kernel: the largest share of a training step's 3x3 time
    kern, (ms_tot, fl_tot, nl, bound_tot) = max(kernels.items(), key=lambda kv: kv[1][0])
    achieved = fl_tot / (ms_tot * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None,
            "kernel": f"{kern} ({nl} launches of a training step, batch {batch}, 128^2)",
            "kernel_ms": round(ms_tot / nl, 4), "kernel_flop": fl_tot / nl,
            "kernel_share": {k: round(v[0] / sum(x[0] for x in kernels.values()), 3) for k, v in kernels.items()},
            # fraction of the attainable: per launch max(FLOP / MFMA peak,
            # activation bytes / HBM peak), summed over the kernel's launches,
            # over their measured time (32-channel levels are HBM-bound:
            # 9 Cin Cout / (Cin + Cout) = 144 FLOP/B < 2.5 PFLOP/s / 8 TB/s)
            "attainable_frac": round(bound_tot / ms_tot, 4),
            "hbm_bound_launches": sum(1 for m, f, n, p, b in launches
                                      if b > f / (MFMA_BF16_PEAK_TFLOPS * 1e9)),
            "train_step_ms": round(step_ms, 3),
            "train_step_TFLOPs": round(step_flop / (step_ms * 1e-3) / 1e12, 1), "forward_ms": round(fwd_ms, 3),
            "forward_TFLOPs": round(fwd_flop / (fwd_ms * 1e-3) / 1e12, 1),
            "launches": {f"{n}/{p}": round(f / (m * 1e-3) / 1e12, 1) for m, f, n, p, b in launches}}


def bench_warmstart(args):
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    from smili_amd.warmstart import data as D
    from smili_amd.warmstart import solve as S
    from smili_amd.warmstart.unet import DEFAULT_WEIGHTS, UNet
    net, meta = UNet.from_file(DEFAULT_WEIGHTS, batch=1)
    base = 2_000_000 + 1000 * args.seed          # held out: training seeds are below 1e6
    seeds = list(range(base, base + args.steps))
    if args.ws_child >= 0:
        # one of --workers processes: its share of the seeds, one JSON row each
        for s in range(base - args.warmup, base):
            S.time_to_image(net, s)
        for s in seeds[args.ws_child::args.workers]:
            r = S.time_to_image(net, s)
            print("ROW " + json.dumps({k: (float(v) if isinstance(v, (np.floating, np.integer)) else v)
                                       for k, v in r.items() if np.isscalar(v) or v is None}), flush=True)
        net.close()
        return
    rows = []
    if args.workers > 1:
        # K processes share the GPU and the host's cores; each clocks its own
        # plain and warm solves, so both clocks see the same contention
        import subprocess
        procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--workload", "warmstart",
                                   "--steps", str(args.steps), "--warmup", str(args.warmup), "--seed", str(args.seed),
                                   "--workers", str(args.workers), "--ws-child", str(i), "--no-cpu"],
                                  stdout=subprocess.PIPE, text=True) for i in range(args.workers)]
        for p in procs:
            for ln in p.stdout:
                if ln.startswith("ROW "):
                    rows.append(json.loads(ln[4:]))
                    if len(rows) % 50 == 0:
                        print(f"[warmstart] {len(rows)}/{args.steps} seeds", file=sys.stderr, flush=True)
        if any(p.wait() != 0 for p in procs):
            raise SystemExit("a warm-start worker failed")
        rows.sort(key=lambda r: r["seed"])
    else:
        for s in range(base - args.warmup, base):
            S.time_to_image(net, s)
        for s in seeds:
            rows.append(S.time_to_image(net, s))
            r = rows[-1]
            # progress on stderr (stdout carries the one JSON line)
            print(f"[warmstart] seed {s}: plain {r['plain_s']:.3f} s, warm "
                  f"{r['warm_s'] if r['warm_s'] is None else round(r['warm_s'], 3)} s", file=sys.stderr, flush=True)
    sm = S.summary(rows)
    roof = _unet_roofline(net.params)
    cpu = None
    if not args.no_cpu:
        case = D.make_case(base)
        prep = D.prepare_case(case)
        per, n, el, thr = _cpu_time_per_eval(dict(prep.kwargs, x0=prep.x0), args.cpu_budget)
        nfev = float(np.median([r["plain_iters"] for r in rows])) * 1.05
        cpu = {"value": round(per * nfev, 3), "unit": "s per image (plain L-BFGS-B to the matched objective)",
               "cores": thr, "kind": "port",
               "sample": f"{n} evaluations of one 128^2 case ({el:.1f} s, C + OpenMP FP64 restatement of the "
                         f"cost) times the median plain evaluation count to the matched objective"}
    net.close()
    line = {
        "metric": "time-to-image speedup, learned warm start vs plain L-BFGS-B (128^2, matched objective 1e-6)",
        "value": round(sm["median_time_ratio"], 3) if sm["median_time_ratio"] else None,
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * sm["median_warm_s"], 2) if sm["median_warm_s"] else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network (f32 accumulation) + f64 imaging cost",
        "data": "synthetic",
        "config": {"workload": "warmstart: held-out seeds of the training distribution (rings, crescents, 1-5 "
                               "Gaussians, disks; 0.3-2 Jy), 128x128 at 1 uas, 6-10 stations at 230 GHz, "
                               "full-visibility chi-square + L1 + TSV, L-BFGS-B m=10",
                   "seeds": [base, base + args.steps], "reached": sm["reached"], "n": sm["n"],
                   "median_iter_ratio": sm["median_iter_ratio"], "median_plain_s": sm["median_plain_s"],
                   "median_warm_s": sm["median_warm_s"], "max_nrmse": sm["max_nrmse"],
                   "proposals_used": sm["proposals_used"],
                   "n_train": int(meta.get("n_seen", meta.get("n_train", 0))),
                   "training_parts": int(meta.get("parts", 1)), "workers": args.workers},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_warmstart_movie(args):
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    from smili_amd import workloads
    from smili_amd.cost import MovieProblem
    from smili_amd.warmstart import movie as WM
    from smili_amd.warmstart.unet import DEFAULT_WEIGHTS, UNet
    net, meta = UNet.from_file(DEFAULT_WEIGHTS, batch=16)
    rows, prep = [], None
    for k in range(args.steps):
        prep, vm = workloads.dynamic_movie(seed=5 + args.seed + k)
        if k == 0:
            with MovieProblem(**prep.kwargs) as g:          # untimed: library and code objects paged in
                g.cost_grad(prep.x0)
                _timed_loop(g, args.warmup, 0)
        rows.append(WM.time_to_image_movie(net, prep, vm["vis"], 2.0, truth=vm["frames"]))
        r = rows[-1]
        print(f"[warmstart_movie] movie seed {5 + args.seed + k}: plain {r['plain_s']:.2f} s "
              f"({r['plain_iters']} it), warm {r['warm_s']} s ({r['warm_iters']} it)",
              file=sys.stderr, flush=True)
    # the GPU side of every iteration: the movie cost's dominant kernel
    kw = dict(prep.kwargs)
    with MovieProblem(**kw, tune=dict(precision=64)) as g:
        g.cost_grad(prep.x0)
        n_ev = 100
        ms = _timed_loop(g, n_ev, 5) / n_ev * 1e3
        roof = _roofline(g, [dict(kw, units=_movie_units(kw))], True, ms, n_ev, "movie")
    from smili_amd.warmstart.solve import time_ratio
    ratios = [time_ratio(r) for r in rows]          # unreached warm solves count as losses
    cpu = None
    if not args.no_cpu:
        c = cpu_baseline_movie(prep, args.cpu_budget)
        nfev = float(np.median([r["plain_iters"] * r["plain_nfev"] / max(r["plain_total_iters"], 1)
                                for r in rows]))
        cpu = {"value": round(nfev / c["value"], 1), "unit": "s per movie (plain L-BFGS-B to the matched objective)",
               "cores": c["cores"], "kind": "port",
               "sample": c["sample"] + f"; times the plain solve's {nfev:.0f} evaluations to the matched objective"}
    net.close()
    med = lambda key: float(np.median([r[key] for r in rows if r[key] is not None])) if rows else None
    line = {
        "metric": "movie time-to-image speedup, per-frame learned warm start vs plain L-BFGS-B "
                  "(64 x 128^2, matched objective 1e-6)",
        "value": round(float(np.median(ratios)), 3) if ratios else None,
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * med("warm_s"), 1) if ratios else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network (f32 accumulation) + f64 imaging cost",
        "data": "synthetic",
        "config": {"workload": "dynamic_movie: 64 frames of 128x128 at 2 uas (fine grids 256x256), 3e5 complex "
                               "visibilities (2e3..8e3 per frame, 10 stations, 12 h), rotating-asymmetry ring, "
                               "per-frame chi-square + L1 + TSV + Rt as one joint cost, L-BFGS-B m=10; the "
                               "network (128x128 at 1 uas) proposes every frame, 2x2-binned onto the movie grid",
                   "movie_seeds": [5 + args.seed, 5 + args.seed + args.steps],
                   "rows": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for r in rows],
                   "median_plain_s": med("plain_s"), "median_warm_s": med("warm_s"),
                   "reached": sum(bool(r["reached"]) for r in rows), "n": len(rows),
                   "n_train": int(meta.get("n_seen", meta.get("n_train", 0)))},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true",
                    help="replay a captured hipGraph instead of launching the kernels one by one")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp64",
                    help="NUFFT arithmetic (chi-square and regularizers are FP64 in both)")
    ap.add_argument("--workload", choices=("single_large", "movie", "survey", "warmstart", "warmstart_movie"),
                    default="single_large")
    ap.add_argument("--workers", type=int, default=1,
                    help="warmstart: processes sharing the GPU, each timing its share of the seeds")
    ap.add_argument("--ws-child", type=int, default=-1, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-run", action="store_true",
                    help="only evaluations plus the calibration copy (run under rocprofv3 --pmc)")
    ap.add_argument("--pmc-ingest", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn two rocprofv3 --pmc counter files into profiles/traffic_*.json")
    args = ap.parse_args()
    if args.pmc_ingest:
        ingest_pmc(*args.pmc_ingest, traffic_file(args.workload, args.precision), args.workload, args.precision)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the objective+gradient path has no multi-GPU "
              "component in this build (it would shard uv points by timestamp and all-reduce the "
              "gradient); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "warmstart":
        if "--steps" not in sys.argv:
            args.steps = 20
        if "--warmup" not in sys.argv:
            args.warmup = 2
        bench_warmstart(args)
        return
    if args.workload == "warmstart_movie":
        if "--steps" not in sys.argv:
            args.steps = 1
        if "--warmup" not in sys.argv:
            args.warmup = 5
        bench_warmstart_movie(args)
        return

    import numpy as np
    import torch

    torch.cuda.set_device(0)
    from smili_amd import workloads
    from smili_amd.cost import BatchProblem, CostProblem, MovieProblem

    fp64 = args.precision == "fp64"
    tune = dict(use_graph=1 if args.graph else 0, precision=64 if fp64 else 32)
    t_setup = time.perf_counter()
    if args.workload == "single_large":
        prep, _ = workloads.single_large(seed=args.seed)
        kw = prep.kwargs
        g = CostProblem(**kw, tune=tune)
        x0 = prep.x0
        kws = [kw]
        per_step = 1
        metric, unit = "obj+grad evals/s (256^2, 1e5 vis)", "evals/s"
        desc = ("single_large: 256x256 (fine grid 512x512, ES w=7, eps 1e-6), 1e5 amplitudes in a "
                "10 Glambda disk + 110,000 closure phases + 495,000 log closure amplitudes "
                "(12 stations x 500 timestamps), L1+TV+TSV+KL")
        extra = {"npix": int(np.asarray(kw["xidx"]).size)}
    elif args.workload == "movie":
        prep, meta = workloads.dynamic_movie(seed=5 + args.seed)
        kw = dict(prep.kwargs)
        g = MovieProblem(**kw, tune=tune)
        x0 = prep.x0
        kws = [dict(kw, units=_movie_units(kw))]
        per_step = 1
        metric, unit = "movie obj+grad evals/s (64 x 128^2, 3e5 vis)", "evals/s"
        desc = ("dynamic_movie: 64 frames of 128x128 (fine grids 256x256), 3e5 complex visibilities "
                "(2e3..8e3 per frame, 10 stations, 12 h), rotating-asymmetry ring, per-frame chi-square "
                "+ L1 + TSV + Rt as one joint cost")
        extra = {"frames": int(prep.Nt), "npix_total": int(x0.size),
                 "vis_per_frame": [int(meta["counts"].min()), int(meta["counts"].max())]}
    else:
        preps, meta = workloads.survey(seed=3 + args.seed)
        kws = [p.kwargs for p in preps]
        g = BatchProblem(kws, tune=tune)
        x0 = np.concatenate([p.x0 for p in preps])
        per_step = len(preps)
        metric, unit = "survey problem-evals/s (512 problems, 64^2..256^2)", "problem-evals/s"
        desc = ("survey: 512 independent problems, sizes uniform over {64,96,128,192,256}^2, "
                "2e3..5e4 complex visibilities (log-uniform) from 6..12 stations, L1 x TSV log grid "
                "8 x 8, TV and MEM on/off; one step evaluates all 512")
        extra = {"problems": len(preps), "nvis_total": int(sum(m["nvis"] for m in meta)),
                 "npix_total": int(x0.size)}
    info = g.info()
    t_setup = time.perf_counter() - t_setup
    if args.pmc_run:
        g.cost_grad(x0)
        dt = _timed_loop(g, args.steps, args.warmup)
        src = torch.ones(COPY_CAL_BYTES // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_run": args.workload, "steps": args.steps, "ms_per_step": dt / args.steps * 1e3}),
              flush=True)
        return
    # x goes to the handle's own device buffer once; every timed step is one
    # evaluation on resident data (the kernels launched one by one, or the
    # captured graph of all of them with --graph)
    first = g.cost_grad(x0)
    dt = _timed_loop(g, args.steps, args.warmup)
    ms = dt / args.steps * 1e3
    value = per_step * args.steps / dt
    again = g.cost_grad(x0)
    assert np.array_equal(np.asarray(first[0]), np.asarray(again[0]))
    info = g.info()
    # host-pointer API (x host->device, grad device->host each call): PCIe-inclusive
    # host-pointer API (x host->device, grad device->host each call): the
    # imager's path works in the handle's page-locked x and grad
    # (host_buffers), a caller with its own arrays pays two host copies more
    nh = max(20, min(args.steps // 5, 200))
    if per_step == 1:
        hx, hg = g.host_buffers()
        hx[:] = x0
        g.cost_grad(hx, out=hg)
        th = time.perf_counter()
        for _ in range(nh):
            g.cost_grad(hx, out=hg)
        host_rate = nh / (time.perf_counter() - th)
    else:
        host_rate = None
    th = time.perf_counter()
    for _ in range(nh):
        g.cost_grad(x0)
    host_rate_copies = per_step * nh / (time.perf_counter() - th)
    roofline = _roofline(g, kws, fp64, ms, args.steps, args.workload)

    cpu = None
    if not args.no_cpu:
        if args.workload == "single_large":
            cpu = cpu_baseline(prep, args.cpu_budget)
        elif args.workload == "movie":
            cpu = cpu_baseline_movie(prep, args.cpu_budget)
        else:
            cpu = cpu_baseline_survey(preps, args.cpu_budget)

    cost = first[0] if per_step == 1 else float(np.sum(first[0]))
    line = {
        "metric": metric,
        "value": round(value, 2),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if fp64 else "f32") + " NUFFT + f64 chi-square/regularizers",
        "data": "synthetic",
        "config": dict({"workload": desc, "nuv": info["nuv"], "graph": bool(info["graph"]),
                        "setup_s": round(t_setup, 2), "precision": args.precision, "cost": cost,
                        "host_api_" + unit.replace("/", "_"): None if host_rate is None else round(host_rate, 1),
                        "host_api_own_arrays_" + unit.replace("/", "_"): round(host_rate_copies, 1)}, **extra),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
