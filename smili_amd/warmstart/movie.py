"""The learned warm start applied per frame of a movie (dynamic imaging,
the joint objective of fftim3d.calc_cost, src/fortlib/fftim3d.f90:472-766).

Each frame's dirty image and dirty beam, made from that frame's own
visibilities on the grid the network was trained on (``data.CFG``: 128 x
128 at 1 uas), go through the network in one batch.  Every proposal is
normalised to the flux of the frame's prior, binned onto the movie's grid
(an integer multiple of the network's pixel, both grids centred on their
reference pixel), and the stacked frames get one scale fitted along the
movie's objective (the secant of ``solve.propose``).  The joint L-BFGS-B
(``imaging.lbfgs.solve``) then runs from it; the movie's prior frames stay
the fallback when the proposal's objective is not lower.
"""
from __future__ import annotations

import time
from types import SimpleNamespace

import numpy as np

from . import data as D


def frame_inputs(vis, nz, cfg=D.CFG):
    """Network inputs [nz, n, n, 2] (dirty image / peak, dirty beam) of the
    frames of a visibility table with a ``frmidx`` column."""
    out = []
    for z in range(nz):
        sel = np.asarray(vis["frmidx"]) == z
        case = SimpleNamespace(u=np.asarray(vis["u"])[sel], v=np.asarray(vis["v"])[sel],
                               vis=np.asarray(vis["amp"])[sel] * np.exp(1j * np.radians(np.asarray(vis["phase"])[sel])),
                               sigma=np.asarray(vis["sigma"])[sel])
        if case.u.size == 0:
            out.append(np.zeros((cfg["n"], cfg["n"], 2), np.float32))
            continue
        dimg, beam, _ = D.dirty_maps(case, cfg)
        out.append(np.stack([dimg, beam], axis=-1))
    return np.stack(out)


def bin_to_grid(img, n_out, factor):
    """Sum factor x factor blocks of a network image [n, n] onto an n_out x
    n_out grid of pixels factor times larger; both grids have their
    reference pixel at (n + 1) / 2, so the blocks are centred (n must be
    even and n / factor <= n_out)."""
    n = img.shape[0]
    if n % factor or (n_out - n // factor) % 2:
        raise ValueError(f"cannot bin a {n}-pixel grid by {factor} onto {n_out} pixels")
    small = img.reshape(n // factor, factor, n // factor, factor).sum(axis=(1, 3))
    out = np.zeros((n_out, n_out), img.dtype)
    o = (n_out - n // factor) // 2
    out[o:o + n // factor, o:o + n // factor] = small
    return out


def propose_movie(net, vis, prep, prob, pix_uas, cfg=D.CFG):
    """(x0 of the movie in scaled units or None, timing dict).  A frame
    whose proposal is empty (no visibilities, or a network image without
    flux inside the movie's window) keeps its prior frame; with no proposal
    at all the result is None."""
    t0 = time.perf_counter()
    nz, n_m = prep.Nt, prep.Nx
    factor = int(round(pix_uas / cfg["pix_uas"]))
    if factor < 1 or abs(factor * cfg["pix_uas"] - pix_uas) > 1e-9 * pix_uas or prep.Nx != prep.Ny:
        raise ValueError("the movie's pixel must be an integer multiple of the network's on a square grid")
    x_in = frame_inputs(vis, nz, cfg)
    t1 = time.perf_counter()
    shapes = net.forward(x_in).astype(np.float64)
    t2 = time.perf_counter()
    npix = prep.xidx.size
    prior = prep.x0.reshape(nz, npix)
    s = np.array(prior, dtype=np.float64).reshape(-1)
    used = 0
    for z in range(nz):
        img = bin_to_grid(shapes[z], n_m, factor)[prep.yidx - 1, prep.xidx - 1]
        if img.sum() > 0 and x_in[z].any():
            s[z * npix:(z + 1) * npix] = img * (prior[z].sum() / img.sum())
            used += 1
    if used == 0:
        return None, dict(dirty_s=t1 - t0, net_s=t2 - t1, fit_s=0.0, frames_proposed=0)
    a1, a2 = 1.0, 1.25
    g1 = float(prob.cost_grad(a1 * s)[1] @ s)
    g2 = float(prob.cost_grad(a2 * s)[1] @ s)
    a = a1 - g1 * (a2 - a1) / (g2 - g1) if g2 != g1 else a1
    a = float(np.clip(a, 0.2, 5.0))
    return a * s, dict(dirty_s=t1 - t0, net_s=t2 - t1, fit_s=time.perf_counter() - t2, frames_proposed=used)


def warm_solve_movie(net, vis, prep, pix_uas, niter=1000, m=10, stop_cost=None, tune=None):
    """Warm-started joint solve: (x scaled frame-major, info dict with the
    SolveInfo, the start used and the objectives of the prior and the
    proposal)."""
    from ..cost import MovieProblem
    from ..imaging import lbfgs
    t0 = time.perf_counter()
    with MovieProblem(**prep.kwargs, tune=tune) as prob:
        x0, tm = propose_movie(net, vis, prep, prob, pix_uas)
        f_prior = prob.cost_grad(prep.x0)[0]
        f_prop = prob.cost_grad(x0)[0] if x0 is not None else float("inf")
        start = "proposal"
        if not f_prop < f_prior:
            x0, start = prep.x0, "prior"
        t_setup = time.perf_counter() - t0
        x, info = lbfgs.solve(prob, x0, niter=niter, m=m, verbose=False, stop_cost=stop_cost)
    return x, dict(info=info, start=start, f_prior=f_prior, f_proposal=f_prop, setup_s=t_setup, **tm)


def time_to_image_movie(net, prep, vis, pix_uas, rel=1e-6, niter=1000, m=10, tune=None, truth=None,
                        track=False):
    """Plain joint solve from the movie's prior frames, then the warm-started
    one stopped at the plain solve's final objective f* (1 + rel).  Times
    are to the first iterate at or below that target: the plain one counts
    the handle's set-up and the solve up to that iterate, the warm one
    everything from the handle's set-up on (dirty maps, network, scale
    fit, both start objectives, the solve).  ``plain_curve`` /
    ``warm_curve``: the iterations at which each solve first reaches
    f* (1 + gap) for the gaps of ``solve.GAPS`` (None: not reached).

    Image distances (NRMSE over all frames): the warm movie against the
    plain final one; with ``track``, also the plain solve's own iterate at
    the target against its final one (how far apart two movies at the same
    objective lie along the plain path; it copies every plain iterate, which
    adds to the plain clock, so timing runs leave it off); with ``truth``
    ([Nt, Ny, Nx] in Jy/pixel), both movies against it."""
    from ..cost import MovieProblem
    from ..imaging import lbfgs
    from .solve import GAPS, _reach
    kept = []                      # plain iterates within rel of the latest cost

    def keep(it, xk, terms):
        c = terms["cost"]
        kept[:] = [(k, ck, xc) for k, ck, xc in kept if ck <= c * (1.0 + rel)]
        kept.append((it, c, np.array(xk, copy=True)))

    t0 = time.perf_counter()
    with MovieProblem(**prep.kwargs, tune=tune) as prob:
        t_setup = time.perf_counter() - t0
        xp, ip = lbfgs.solve(prob, prep.x0, niter=niter, m=m, verbose=False, callback=keep if track else None)
    target = ip.cost * (1.0 + rel)
    kp, tp = _reach(ip, target)
    x_tgt = next((xc for k, ck, xc in kept if ck <= target), xp)
    t0 = time.perf_counter()
    xw, w = warm_solve_movie(net, vis, prep, pix_uas, niter=niter, m=m, stop_cost=target, tune=tune)
    wall_w = time.perf_counter() - t0
    iw = w["info"]
    kw, tw = _reach(iw, target)

    def dist(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    out = dict(frames=int(prep.Nt), npix=int(xp.size), gap_prior=w["f_prior"] / ip.cost - 1.0,
               gap_proposal=w["f_proposal"] / ip.cost - 1.0, start=w["start"],
               plain_iters=kp, plain_s=t_setup + tp, plain_total_iters=ip.niter, plain_nfev=ip.nfev,
               plain_total_s=t_setup + ip.wall_s, plain_cost=ip.cost,
               warm_iters=kw, warm_nfev=iw.nfev, warm_s=(w["setup_s"] + tw) if kw is not None else None,
               warm_wall_s=wall_w,
               warm_cost=iw.cost, reached=kw is not None, nrmse=dist(xw, xp),
               nrmse_plain_at_target=dist(x_tgt, xp) if track else None,
               dirty_s=w["dirty_s"], net_s=w["net_s"], fit_s=w["fit_s"], frames_proposed=w["frames_proposed"],
               gaps=list(GAPS), plain_curve=[_reach(ip, ip.cost * (1.0 + g))[0] for g in GAPS],
               warm_curve=[_reach(iw, ip.cost * (1.0 + g))[0] for g in GAPS])
    if truth is not None:
        tv = np.asarray(truth)[:, prep.yidx - 1, prep.xidx - 1].reshape(-1)
        out["truth_err_plain"] = dist(xp / prep.inormfactr, tv)
        out["truth_err_warm"] = dist(xw / prep.inormfactr, tv)
    return out
