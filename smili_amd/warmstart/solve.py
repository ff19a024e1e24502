"""Warm-started imaging and the time-to-image comparison.

``propose``: network image from the dirty image and beam, scaled by the
factor that minimises the imaging objective along it (the objective is
close to quadratic in a scale factor: chi-square and TSV are quadratic,
the smoothed L1 linear for positive pixels, so a secant through two
directional derivatives lands on the minimiser).

``warm_solve``: the plain L-BFGS-B imager (smili_amd.imaging.lbfgs.solve,
the reference's src/fortlib/fftim2d.f90:286-351 loop) from the proposal;
the prior image stays the fallback when the proposal's objective is not
lower than the prior's.

``time_to_image``: for one seed, the plain solve from the flat prior and
the warm-started solve, each timed from the visibility table to the first
iterate whose objective is within ``rel`` (1e-6) of the plain solve's
final objective.  Both clocks include the same set-up (prepare, handle
creation); the warm start adds its dirty maps, network and scale fit.
"""
from __future__ import annotations

import time

import numpy as np

from . import data as D


def _vec(img, prep):
    return np.asarray(img, np.float64)[prep.yidx - 1, prep.xidx - 1] * prep.inormfactr


def _img(x, prep):
    out = np.zeros((prep.Ny, prep.Nx))
    out[prep.yidx - 1, prep.xidx - 1] = x / prep.inormfactr
    return out


def propose(net, case, prep, prob, cfg=D.CFG):
    """(x0 scaled vector of the proposal, timing dict)."""
    t0 = time.perf_counter()
    dimg, beam, _ = D.dirty_maps(case, cfg)
    t1 = time.perf_counter()
    shape = net.forward(np.stack([dimg, beam], axis=-1)[None])[0].astype(np.float64)
    t2 = time.perf_counter()
    s = _vec(shape, prep)
    if not np.any(s > 0):
        return None, dict(dirty_s=t1 - t0, net_s=t2 - t1, fit_s=0.0)
    a1 = case.flux * prep.inormfactr / s.sum()
    a2 = 1.25 * a1
    g1 = float(prob.cost_grad(a1 * s)[1] @ s)
    g2 = float(prob.cost_grad(a2 * s)[1] @ s)
    a = a1 - g1 * (a2 - a1) / (g2 - g1) if g2 != g1 else a1
    a = float(np.clip(a, 0.2 * a1, 5.0 * a1))
    return a * s, dict(dirty_s=t1 - t0, net_s=t2 - t1, fit_s=time.perf_counter() - t2)


def warm_solve(net, case, cfg=D.CFG, stop_cost=None, tune=None):
    """Warm-started imaging of one case: (image [n, n] Jy/pixel, info dict)."""
    from ..cost import CostProblem
    from ..imaging import lbfgs
    t0 = time.perf_counter()
    prep = D.prepare_case(case, cfg)
    with CostProblem(**prep.kwargs, tune=tune) as prob:
        x0, tm = propose(net, case, prep, prob, cfg)
        f_prior = prob.cost_grad(prep.x0)[0]
        f_prop = prob.cost_grad(x0)[0] if x0 is not None else float("inf")
        used = "proposal"
        if f_prop >= f_prior:
            x0, used = prep.x0, "prior"
        t_setup = time.perf_counter() - t0
        x, info = lbfgs.solve(prob, x0, niter=cfg["niter"], m=cfg["m"], verbose=False, stop_cost=stop_cost)
    return _img(x, prep), dict(info=info, setup_s=t_setup, start=used, f_prior=f_prior, f_proposal=f_prop, **tm)


def _reach(info, target):
    """(iterations, seconds) of the first iterate at or below target."""
    for k, (c, t) in enumerate(zip(info.history, info.history_t)):
        if c <= target:
            return k + 1, t
    return None, None


def time_to_image(net, seed, cfg=D.CFG, rel=1e-6):
    from ..cost import CostProblem
    from ..imaging import lbfgs
    case = D.make_case(seed, cfg)
    # plain solve from the flat prior
    t0 = time.perf_counter()
    prep = D.prepare_case(case, cfg)
    with CostProblem(**prep.kwargs) as prob:
        t_setup = time.perf_counter() - t0
        xp, ip = lbfgs.solve(prob, prep.x0, niter=cfg["niter"], m=cfg["m"], verbose=False)
    target = ip.cost * (1.0 + rel)
    kp, tp = _reach(ip, target)
    img_p = _img(xp, prep)
    # warm start, stopped at the plain solve's final objective
    t0 = time.perf_counter()
    img_w, w = warm_solve(net, case, cfg, stop_cost=target)
    wall_w = time.perf_counter() - t0
    iw = w["info"]
    kw, tw = _reach(iw, target)
    nrmse = float(np.linalg.norm(img_w - img_p) / max(np.linalg.norm(img_p), 1e-300))
    return dict(seed=int(seed), kind=case.kind, nvis=int(case.u.size),
                gap_prior=w["f_prior"] / ip.cost - 1.0, gap_proposal=w["f_proposal"] / ip.cost - 1.0,
                plain_iters=kp, plain_s=t_setup + tp,
                plain_total_iters=ip.niter, plain_cost=ip.cost, warm_iters=kw,
                warm_s=(w["setup_s"] + tw) if tw is not None else None, warm_wall_s=wall_w, warm_cost=iw.cost,
                warm_start=w["start"], reached=kw is not None, nrmse=nrmse,
                net_s=w["net_s"], dirty_s=w["dirty_s"], fit_s=w["fit_s"])


GAPS = (10.0, 1.0, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6)


def gap_curve(net, seed, cfg=D.CFG, gaps=GAPS):
    """Iterations at which the plain solve (from the flat prior) and the
    warm-started solve first reach objective f* (1 + gap), f* the plain
    solve's final objective: {"plain": [...], "warm": [...]} per gap (None:
    not reached)."""
    from ..cost import CostProblem
    from ..imaging import lbfgs
    case = D.make_case(seed, cfg)
    prep = D.prepare_case(case, cfg)
    with CostProblem(**prep.kwargs) as prob:
        _, ip = lbfgs.solve(prob, prep.x0, niter=cfg["niter"], m=cfg["m"], verbose=False)
    fstar = ip.cost
    _, w = warm_solve(net, case, cfg, stop_cost=fstar * (1.0 + min(gaps)))

    def first(hist, g):
        return next((k + 1 for k, c in enumerate(hist) if c <= fstar * (1.0 + g)), None)
    return dict(seed=int(seed), plain=[first(ip.history, g) for g in gaps],
                warm=[first(w["info"].history, g) for g in gaps])


def time_ratio(r):
    """plain / warm time-to-image of one row; a warm solve that never
    reached the matched objective counts as a loss, plain time over its
    whole wall time, capped at 1."""
    if r["reached"]:
        return r["plain_s"] / r["warm_s"]
    return min(r["plain_s"] / r["warm_wall_s"], 1.0)


def summary(rows):
    """Medians over all rows (unreached warm solves count as losses, see
    time_ratio); ``reached`` of ``n`` rows reached the matched objective."""
    ok = [r for r in rows if r["reached"]]
    it_ratio = [r["plain_iters"] / r["warm_iters"] for r in ok]
    t_ratio = [time_ratio(r) for r in rows]
    return dict(n=len(rows), reached=len(ok), median_iter_ratio=float(np.median(it_ratio)) if ok else None,
                median_time_ratio=float(np.median(t_ratio)) if rows else None,
                geomean_time_ratio=float(np.exp(np.mean(np.log(t_ratio)))) if rows else None,
                median_plain_s=float(np.median([r["plain_s"] for r in rows])),
                median_warm_s=float(np.median([r["warm_s"] for r in ok])) if ok else None,
                max_nrmse=float(max(r["nrmse"] for r in ok)) if ok else None,
                median_gap_prior=float(np.median([r["gap_prior"] for r in rows])),
                median_gap_proposal=float(np.median([r["gap_proposal"] for r in rows])),
                proposals_used=sum(r["warm_start"] == "proposal" for r in rows))


def main(argv=None):
    import argparse
    import json
    from .unet import UNet
    ap = argparse.ArgumentParser(description="time-to-image of the learned warm start vs the plain solve")
    ap.add_argument("--weights", required=True)
    ap.add_argument("--seeds", default="1000000:1000020")
    ap.add_argument("--out", default=None)
    ap.add_argument("--gap-curve", action="store_true",
                    help="iterations to each objective gap (plain vs warm) instead of the time-to-image rows")
    a = ap.parse_args(argv)
    net, _ = UNet.from_file(a.weights, batch=1)
    lo, hi = (int(s) for s in a.seeds.split(":"))
    if a.gap_curve:
        curves = []
        for s in range(lo, hi):
            curves.append(gap_curve(net, s))
            print(json.dumps(curves[-1]), flush=True)
        med = {k: [float(np.nanmedian([np.nan if c[k][i] is None else c[k][i] for c in curves]))
                   for i in range(len(GAPS))] for k in ("plain", "warm")}
        print(json.dumps({"gaps": list(GAPS), "plain_median": med["plain"], "warm_median": med["warm"]}), flush=True)
        return 0
    rows = []
    for s in range(lo, hi):
        r = time_to_image(net, s)
        rows.append(r)
        print(json.dumps({k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}), flush=True)
    sm = summary(rows)
    print(json.dumps(sm), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump({"rows": rows, "summary": sm}, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
