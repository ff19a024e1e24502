"""Static imager with SMILI's API: ``imaging(initimage, ..., vistable=...,
l1_lambda=..., ...)`` (smili/imaging/lbfgs.py:44-576).

The objective+gradient runs on the GPU (``CostProblem``, the C ABI of
libsmili_hip.so) and the optimiser is L-BFGS-B (Byrd, Lu, Nocedal & Zhu,
v3.0 of Zhu, Byrd, Lu & Nocedal / Morales & Nocedal 2011 -- the algorithm of
the reference's src/lbfgsb/lbfgsb.f) in FP64 on the host: the package's
native, threaded implementation (``smili_amd.lbfgsb``, libsmili_lbfgsb.so)
by default, or SciPy's implementation of the same algorithm.  The reference's reverse-communication
loop (src/fortlib/fftim2d.f90:286-351) stops once the iteration counter
exceeds ``niter`` at a NEW_X, i.e. after ``niter + 1`` iterations; ``m``,
``factr`` and ``pgtol`` come from the module dict ``lbfgsbprms``
(lbfgs.py:34-38).
"""
from __future__ import annotations

import copy
import time
from dataclasses import dataclass, field

import numpy as np
from scipy.optimize import Bounds, minimize

from .. import lbfgsb as LB
from ..cost import CostProblem
from . import deburr as D
from .prep import prepare

lbfgsbprms = {
    "m": 5,
    "factr": 1e1,
    "pgtol": 0.,
}


@dataclass
class SolveInfo:
    """Optimiser outcome (the reference prints it only, fftim2d.f90:377)."""
    niter: int
    nfev: int
    cost: float
    terms: dict
    message: str
    success: bool
    wall_s: float
    eval_s: float
    history: list = field(default_factory=list)      # cost after each iteration
    history_t: list = field(default_factory=list)    # seconds since the start, after each iteration
    stopped: bool = False                            # reached stop_cost
    start: str = "prior"                             # "proposal" when a warm start was used


def _summary(terms, inormfactr, x, nz=1):
    """The summary calc_cost prints (fftim2d.f90:557-606; for movies
    fftim3d.f90:649-713)."""
    lines = ["  Cost Function          : %13.6e" % terms["cost"],
             "  Data term              : %13.6e" % terms["chisq"]]
    for k, lab in (("chisq_fcv", "Complex Visibilities"), ("chisq_amp", "Amplitudes"),
                   ("chisq_cp", "Closure Phase"), ("chisq_ca", "Closure Amplitudes")):
        if terms[k] != 0:
            lines.append("    %-20s : %13.6e" % (lab, terms[k]))
    lines.append("  Reguarization term     : %13.6e" % terms["reg"])
    for k, lab in (("l1", "l1-norm"), ("tv", "TV"), ("tsv", "TSV"), ("kl", "KL divergence"),
                   ("gs", "GS entropy"), ("tfd", "Total Flux Density"), ("cen", "Centoroid Reg."),
                   ("sm", "second momentum")):
        if terms[k] != 0:
            lines.append("    %-20s : %13.6e" % (lab, terms[k]))
    if "lc" in terms and terms["lc"] != 0:
        lines.append("    %-20s : %13.6e" % ("Light curve", terms["lc"]))
    if "rt" in terms:
        lines.append("  Total flux (avg)       : %13.6e" % (np.sum(x) / nz / inormfactr))
        for k, lab in (("rt", "Dynamical  Rt Reg."), ("ri", "Dynamical  Ri Reg.")):
            if terms[k] != 0:
                lines.append("  %-22s : %13.6e" % (lab, terms[k]))
        return "\n".join(lines)
    lines.append("  Total flux             : %13.6e" % (np.sum(x) / inormfactr))
    return "\n".join(lines)


def _setulb_loop(fg, x0, nonneg, m, factr, pgtol, maxiter, maxls, callback, xbuf=None):
    """L-BFGS-B through SciPy's own v3.0 routine ``setulb`` and its
    reverse-communication protocol, the loop fftim2d.f90:286-351 runs over
    lbfgsb.f's setulb: task FG -> evaluate f, g at x; task NEW_X -> one
    iteration done (the callback, then the iteration limit); anything else
    ends the loop.  ``scipy.optimize.minimize`` drives the same routine the
    same way, but builds its bound arrays in a Python loop over every
    variable (seconds at a million pixels); here they are set whole.
    ``xbuf``: an array of n float64 to keep the iterate in (the handle's
    page-locked x, so an evaluation reads it without a host copy).
    Returns (x, iterations, message, success) or None when this SciPy's
    ``setulb`` does not take the call (then ``minimize`` is used)."""
    try:
        from scipy.optimize import _lbfgsb
        from scipy.optimize._lbfgsb_py import status_messages, task_messages
    except ImportError:
        return None
    n = x0.size
    if xbuf is not None:
        x = xbuf
        x[:] = np.maximum(x0, 0.0) if nonneg else x0
    else:
        x = np.array(np.maximum(x0, 0.0) if nonneg else x0, dtype=np.float64)
    nbd = np.full(n, 1 if nonneg else 0, np.int32)       # 1: lower bound only
    low, upp = np.zeros(n), np.zeros(n)
    f, g = 0.0, np.zeros(n)
    wa = np.zeros(2 * m * n + 5 * n + 11 * m * m + 8 * m)
    iwa = np.zeros(3 * n, np.int32)
    task, ln_task = np.zeros(2, np.int32), np.zeros(2, np.int32)
    lsave, isave, dsave = np.zeros(4, np.int32), np.zeros(44, np.int32), np.zeros(29)
    nit, first = 0, True
    while True:
        try:
            _lbfgsb.setulb(m, x, low, upp, nbd, f, g, factr, pgtol, wa, iwa, task, lsave, isave, dsave,
                           maxls, ln_task)
        except TypeError:
            if first:
                return None
            raise
        first = False
        if task[0] == 3:                                     # FG
            f, g = fg(x)
            g = np.ascontiguousarray(g, dtype=np.float64)
        elif task[0] == 1:                                   # NEW_X
            nit += 1
            xv = x.view()
            xv.flags.writeable = False
            try:
                callback(xv)
            except StopIteration:
                task[0], task[1] = 5, 505
            if nit >= maxiter:
                task[0], task[1] = 5, 504
        else:
            break
    x = np.array(x, copy=True)               # the iterate outlives the handle's buffer
    if task[0] == 5 and task[1] == 505:
        return x, nit, "`callback` raised `StopIteration`.", False      # minimize's wording
    return x, nit, status_messages[task[0]] + ": " + task_messages[task[1]], bool(task[0] == 4)


def solve(problem: CostProblem, x0, niter=1000, nonneg=True, m=None, factr=None, pgtol=None,
          nprint=500, inormfactr=1.0, callback=None, verbose=True, stop_cost=None, driver="auto"):
    """L-BFGS-B over ``problem.cost_grad`` from x0 (scaled units); with
    ``stop_cost`` it also ends at the first iterate whose cost is at or
    below it.  ``driver``: "native" (libsmili_lbfgsb.so; also "auto"),
    "setulb" (SciPy's routine driven directly), "minimize" (through
    scipy.optimize.minimize; the same iterates as "setulb").
    ``callback(it, xk, terms)`` must not keep xk without copying it.
    Returns (x, SolveInfo)."""
    if driver not in ("auto", "native", "setulb", "minimize"):
        raise ValueError(f'driver must be "auto", "native", "setulb" or "minimize", not {driver!r}')
    m = lbfgsbprms["m"] if m is None else m
    factr = lbfgsbprms["factr"] if factr is None else factr
    pgtol = lbfgsbprms["pgtol"] if pgtol is None else pgtol
    state = {"nfev": 0, "eval_s": 0.0, "it": 0, "last": None, "stopped": False}
    hist, hist_t = [], []
    nz = getattr(problem, "Nz", 1)

    # the handle's page-locked x and grad: setulb keeps its iterate in the
    # first and reads the gradient from the second, so an evaluation copies
    # nothing on the host (the reference's loop passes its x and g to
    # calc_cost in place, fftim2d.f90:286-351)
    bufs = problem.host_buffers() if hasattr(problem, "host_buffers") else None

    def fg(x):
        t = time.perf_counter()
        if bufs is not None:
            c, g, terms = problem.cost_grad(x, out=bufs[1])
        else:
            c, g, terms = problem.cost_grad(x)
        state["eval_s"] += time.perf_counter() - t
        state["nfev"] += 1
        state["last"] = terms
        return c, g

    def cb(xk):
        state["it"] += 1
        hist.append(state["last"]["cost"])
        hist_t.append(time.perf_counter() - t0)
        if verbose and nprint > 0 and state["it"] % nprint == 0:
            print("Iteration :%5d/%5d" % (state["it"], niter))
            print(_summary(state["last"], inormfactr, xk, nz))
        if callback is not None:
            callback(state["it"], xk, state["last"])
        if stop_cost is not None and state["last"]["cost"] <= stop_cost:
            state["stopped"] = True
            raise StopIteration

    x0 = np.asarray(x0, np.float64).ravel()
    t0 = time.perf_counter()
    if verbose:
        c, _, terms = problem.cost_grad(x0)
        print("Summary of the initial image")
        print(_summary(terms, inormfactr, x0, nz))
    if driver in ("auto", "native"):
        x, nit, message, success = LB.minimize(fg, x0, m=int(m), lower=0.0 if nonneg else None, factr=float(factr),
                                               pgtol=float(pgtol), maxiter=int(niter) + 1, maxls=20, callback=cb,
                                               xbuf=None if bufs is None else bufs[0])
        if message == LB.STOPPED_BY_CALLBACK:
            message = "`callback` raised `StopIteration`."      # the wording of the other drivers
        out = x, nit, message, success
    else:
        out = None if driver == "minimize" else _setulb_loop(fg, x0, nonneg, int(m), float(factr), float(pgtol),
                                                              int(niter) + 1, 20, cb,
                                                              None if bufs is None else bufs[0])
    if out is None:
        if driver == "setulb":
            raise RuntimeError("this SciPy's L-BFGS-B routine does not take the direct call")
        def fg_copy(x):
            # minimize caches the gradient it is handed: give it its own array
            c, g = fg(x)
            return c, np.array(g, copy=True)

        res = minimize(fg_copy, x0, jac=True, method="L-BFGS-B",
                       bounds=Bounds(np.zeros(x0.size), np.full(x0.size, np.inf)) if nonneg else None,
                       callback=cb,
                       options=dict(maxcor=int(m), ftol=float(factr) * np.finfo(float).eps, gtol=float(pgtol),
                                    maxiter=int(niter) + 1, maxfun=10 ** 9, maxls=20))
        out = res.x, int(res.nit), str(res.message), bool(res.success)
    x, nit, message, success = out
    wall = time.perf_counter() - t0
    c, _, terms = problem.cost_grad(x)
    if verbose:
        print("Final Summary")
        print("  Iteration ent at %5d/%5d" % (nit, niter))
        print(_summary(terms, inormfactr, x, nz))
        print(message)
    info = SolveInfo(niter=nit, nfev=int(state["nfev"]), cost=float(c), terms=terms,
                     message=message, success=success, wall_s=wall,
                     eval_s=state["eval_s"], history=hist, history_t=hist_t, stopped=state["stopped"])
    return x, info


def _window_of(imregion, initimage, istokes, ifreq):
    if imregion is None:
        return None
    if hasattr(imregion, "imagewin"):
        return imregion.imagewin(initimage, istokes, ifreq)
    if hasattr(imregion, "data"):
        return np.asarray(imregion.data)[0, 0] > 0.5          # lbfgs.py:240-241
    return np.asarray(imregion, bool)


def imaging(initimage, imregion=None, vistable=None, amptable=None, bstable=None, catable=None,
            w_vis=1, w_amp=1, w_bs=1, w_ca=1,
            l1_lambda=-1., l1_prior=None, l1_noise=1e-2,
            tv_lambda=-1, tv_prior=None, tsv_lambda=-1, tsv_prior=None,
            kl_lambda=-1, kl_prior=None, gs_lambda=-1, gs_prior=None,
            tfd_lambda=-1, tfd_tgterror=0.01, cen_lambda=-1, cen_alpha=3,
            sm_lambda=-1, sm_maj=0., sm_min=0., sm_phi=0.,
            pulse_fwhm=-1, niter=1000, nonneg=True, nprint=500, totalflux=None, inorm=1.,
            istokes=0, ifreq=0, tune=None, verbose=True, return_info=False):
    """smili.imaging.lbfgs.imaging on the GPU cost.

    ``initimage`` is an image with ``header`` (nx, ny, dx, dy in degrees,
    nxref, nyref) and ``data[istokes, ifreq, iy, ix]`` (an IMFITS or
    ``smili_amd.imaging.image.Image``); tables are DataFrames or dicts with
    the reference's column names.  ``tune``: smili_tune fields, or "auto"
    (smili_amd.autotune picks them for this problem first).  Priors are images (their data[0, 0] is
    used, lbfgs.py:279-283).  Returns a copy of initimage holding the
    solution (lbfgs.py:565-576); with ``return_info`` also the SolveInfo.
    """
    if isinstance(tune, str) and tune != "auto":
        raise ValueError('tune must be a dict of smili_tune fields, None or "auto"')
    hdr = initimage.header
    window = _window_of(imregion, initimage, istokes, ifreq)

    def prior(p):
        return None if p is None else np.asarray(p.data)[0, 0]

    if pulse_fwhm > 0:
        # the total flux comes from the tables as given (lbfgs.py:203-209),
        # then every table is divided by the pulse's Fourier transform
        # (lbfgs.py:436-478)
        if totalflux is None:
            amps = [np.max(np.asarray(t["amp"], np.float64)) for t in (vistable, amptable) if t is not None]
            totalflux = float(np.max(amps)) if amps else None
        fw = np.deg2rad(pulse_fwhm * _angconv(initimage))
        vistable = None if vistable is None else D.deburr_vistable(vistable, fw)
        amptable = None if amptable is None else D.deburr_vistable(amptable, fw)
        bstable = None if bstable is None else D.deburr_bstable(bstable, fw)
        catable = None if catable is None else D.deburr_catable(catable, fw)
    prep = prepare(np.asarray(initimage.data, np.float64)[istokes, ifreq],
                   np.deg2rad(hdr["dx"]), np.deg2rad(hdr["dy"]), hdr["nxref"], hdr["nyref"], window,
                   vistable, amptable, bstable, catable, w_vis, w_amp, w_bs, w_ca,
                   l1_lambda, prior(l1_prior), l1_noise, tv_lambda, prior(tv_prior),
                   tsv_lambda, prior(tsv_prior), kl_lambda, prior(kl_prior), gs_lambda, prior(gs_prior),
                   tfd_lambda, tfd_tgterror, cen_lambda, cen_alpha, sm_lambda,
                   np.deg2rad(sm_maj * _angconv(initimage)), np.deg2rad(sm_min * _angconv(initimage)), sm_phi,
                   totalflux, inorm)
    if tune == "auto":
        from ..autotune import autotune
        tune, _ = autotune(lambda t: CostProblem(**prep.kwargs, tune=t), prep.x0)
    with CostProblem(**prep.kwargs, tune=tune) as prob:
        x, info = solve(prob, prep.x0, niter=niter, nonneg=nonneg, nprint=nprint,
                        inormfactr=prep.inormfactr, verbose=verbose)
    Iout = x / prep.inormfactr
    out = copy.deepcopy(initimage)
    data = np.array(out.data, dtype=np.float64, copy=True)
    data[istokes, ifreq] = 0.0
    data[istokes, ifreq, prep.yidx - 1, prep.xidx - 1] = Iout
    if pulse_fwhm > 0:
        # the solution convolved with the pulse, total flux kept
        # (lbfgs.py:574-575; IMFITS.convolve_gauss, imdata.py:1705-1774)
        fpx = pulse_fwhm * _angconv(initimage) / abs(hdr["dx"])
        fpy = pulse_fwhm * _angconv(initimage) / abs(hdr["dy"])
        data[istokes, ifreq] = D.convolve_gauss(data[istokes, ifreq], fpx, fpy, save_totalflux=True)
    out.data = data
    if hasattr(out, "update_fits"):
        out.update_fits()
    return (out, info) if return_info else out


def _angconv(image):
    """Degrees per unit of the image's angular unit (sm_maj / sm_min are given
    in that unit, lbfgs.py:413-418)."""
    unit = getattr(image, "angunit", "uas")
    return {"deg": 1.0, "degree": 1.0, "arcmin": 1 / 60, "amin": 1 / 60, "arcsec": 1 / 3600,
            "asec": 1 / 3600, "mas": 1 / 3.6e6, "uas": 1 / 3.6e9}[unit]
