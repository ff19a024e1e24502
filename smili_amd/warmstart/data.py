"""Seeded training and test cases of the learned warm start.

Each case is a synthetic sky (a blurred ring, a crescent, 1-5 Gaussians or
a uniform disk; total flux 0.3-2 Jy), Earth-rotation uv coverage of 6-10
seeded stations at 230 GHz (100 scans over 12 h, 10 deg elevation cut, at
least 400 visibilities),
complex visibilities with thermal noise of 1% of the total flux per real
and imaginary part, and the imaging problem SMILI would solve for it: a
128 x 128 image at 1 micro-arcsecond pixels, full-visibility chi-square +
L1 (lambda 1) + TSV (lambda 1e4), non-negative, from a flat prior (a
circular Gaussian of 20 micro-arcseconds FWHM carrying the total flux).

Network input: the naturally weighted dirty image divided by its peak and
the dirty beam (peak 1), both by the reference's transform convention
(vistable.py:2670-2694).  Network target: the plain GPU imager's solution
(L-BFGS-B, m = 10, up to 1000 iterations) divided by its peak.  The flux
scale is restored at use by a least-squares fit to the visibilities.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from .. import fftlib, synth
from ..synth import UAS

CFG = dict(n=128, pix_uas=1.0, freq=230e9, n_scans=100, hours=12.0, elev_cut=10.0,
           stations=(6, 10), dec=(0.0, 30.0), gst0=(-8.0, -4.0), lon=(-70.0, 70.0), lat=(-30.0, 60.0),
           min_vis=400, noise=0.01, l1=1.0, tsv=1e4, m=10, niter=1000, prior_fwhm_uas=20.0)


@dataclass
class Case:
    seed: int
    kind: str
    flux: float
    u: np.ndarray          # wavelengths
    v: np.ndarray
    vis: np.ndarray        # complex Jy
    sigma: np.ndarray
    sky: list


def geometry(cfg=CFG):
    n = cfg["n"]
    return n, -cfg["pix_uas"] * UAS, cfg["pix_uas"] * UAS, (n + 1) / 2.0


def random_sky(rng):
    kind = ["ring", "crescent", "gaussians", "disk"][int(rng.integers(4))]
    flux = float(rng.uniform(0.3, 2.0))
    off = lambda s: float(rng.uniform(-s, s)) * UAS  # noqa: E731
    if kind == "ring":
        d = rng.uniform(20, 60)
        sky = [synth.Ring(flux, d * UAS, rng.uniform(5, 15) * UAS)]
        if rng.uniform() < 0.5:
            # brightness varying along the ring
            sky = synth.asymmetric_ring(flux, d * UAS, rng.uniform(5, 15) * UAS, rng.uniform(0, 2 * math.pi),
                                        asym=float(rng.uniform(0.2, 0.9)))
    elif kind == "crescent":
        d_out = rng.uniform(25, 60)
        d_in = d_out * rng.uniform(0.4, 0.8)
        a, r = rng.uniform(0, 2 * math.pi), rng.uniform(0.3, 1.0) * (d_out - d_in) / 2
        sky = [synth.Crescent(flux, d_out * UAS, d_in * UAS, r * math.sin(a) * UAS, r * math.cos(a) * UAS,
                              rng.uniform(3, 10) * UAS, off(8), off(8))]
    elif kind == "gaussians":
        k = int(rng.integers(1, 6))
        fl = rng.uniform(0.2, 1.0, k)
        fl *= flux / fl.sum()
        sky = [synth.Gaussian(float(fl[i]), rng.uniform(5, 40) * UAS, off(30), off(30)) for i in range(k)]
    else:
        sky = [synth.Disk(flux, rng.uniform(10, 60) * UAS, rng.uniform(2, 8) * UAS, off(10), off(10))]
    return kind, flux, sky


def make_case(seed, cfg=CFG):
    rng = np.random.default_rng([int(seed), 20241203])
    kind, flux, sky = random_sky(rng)
    for _ in range(100):
        # arrays are redrawn until they give min_vis visibilities
        n_st = int(rng.integers(cfg["stations"][0], cfg["stations"][1] + 1))
        dec = float(rng.uniform(*cfg["dec"]))
        scans = synth.earth_rotation_scans(rng, n_st, dec, cfg["freq"], cfg["n_scans"], cfg["hours"],
                                           elev_cut_deg=cfg["elev_cut"], gst0=float(rng.uniform(*cfg["gst0"])),
                                           lon_deg=cfg["lon"], lat_deg=cfg["lat"])
        if sum(s.u.size for s in scans) >= cfg["min_vis"]:
            break
    u = np.concatenate([s.u for s in scans])
    v = np.concatenate([s.v for s in scans])
    sig = cfg["noise"] * flux
    vis = synth.sky_vis(sky, u, v) + sig * (rng.standard_normal(u.size) + 1j * rng.standard_normal(u.size))
    return Case(int(seed), kind, flux, u, v, vis, np.full(u.size, sig), sky)


def vistable(case):
    """The case as a visibility table (u, v, amp, phase in degrees, sigma)."""
    return {"u": case.u, "v": case.v, "amp": np.abs(case.vis), "phase": np.angle(case.vis, deg=True),
            "sigma": case.sigma}


def dirty_maps(case, cfg=CFG):
    """(dirty image / its peak, dirty beam) as float32 [n, n] (row = y):
    naturally weighted maps on (u, -u) with the reference-pixel phase
    (vistable.py:2670-2694), beam normalised to 1 at the reference pixel.
    Returns also the dirty image's peak in Jy/beam."""
    n, dx, dy, ref = geometry(cfg)
    us = np.concatenate([case.u, -case.u]) * 2 * np.pi * dx * -1
    vs = np.concatenate([case.v, -case.v]) * 2 * np.pi * dy * -1
    w = 1.0 / np.concatenate([case.sigma, case.sigma]) ** 2
    w = w / w.sum()
    ph = np.exp(1j * (us * (n / 2.0 + 1 - ref) + vs * (n / 2.0 + 1 - ref)))
    V = np.concatenate([case.vis, np.conj(case.vis)]) * ph * w
    B = w * ph
    h = fftlib.Nufft(us, vs, n, n)
    try:
        dimg = h.adj_re(V.real, V.imag).reshape(n, n)
        beam = h.adj_re(B.real, B.imag).reshape(n, n)
    finally:
        h.close()
    peak = float(dimg.max())
    return (dimg / peak).astype(np.float32), (beam / beam.max()).astype(np.float32), peak


def prior_image(case, cfg=CFG):
    n, dx, dy, ref = geometry(cfg)
    return synth.sky_image([synth.Gaussian(case.flux, cfg["prior_fwhm_uas"] * UAS)], n, n, dx, dy, ref, ref)


def prepare_case(case, cfg=CFG):
    from ..imaging.prep import prepare
    n, dx, dy, ref = geometry(cfg)
    return prepare(prior_image(case, cfg), dx, dy, ref, ref, vistable=vistable(case),
                   l1_lambda=cfg["l1"], tsv_lambda=cfg["tsv"])


def plain_solve(case, cfg=CFG, tune=None):
    """The plain imager from the flat prior: (image [n, n] Jy/pixel, info)."""
    from ..cost import CostProblem
    from ..imaging import lbfgs
    prep = prepare_case(case, cfg)
    with CostProblem(**prep.kwargs, tune=tune) as prob:
        x, info = lbfgs.solve(prob, prep.x0, niter=cfg["niter"], m=cfg["m"], verbose=False)
    img = np.zeros(prep.Ny * prep.Nx)
    img[(prep.yidx - 1) * prep.Nx + (prep.xidx - 1)] = x / prep.inormfactr
    return img.reshape(prep.Ny, prep.Nx), info


def sample(seed, cfg=CFG, log=None):
    """One training sample: (input [n, n, 2] float32, target [n, n] float32,
    meta dict); ``log(msg)`` receives the phases."""
    t0 = time.perf_counter()
    case = make_case(seed, cfg)
    if log:
        log(f"# {seed} case {time.perf_counter() - t0:.2f} s")
    dimg, beam, peak = dirty_maps(case, cfg)
    if log:
        log(f"# {seed} dirty maps {time.perf_counter() - t0:.2f} s")
    img, info = plain_solve(case, cfg)
    x = np.stack([dimg, beam], axis=-1)
    tmax = float(img.max())
    tgt = (img / tmax).astype(np.float32) if tmax > 0 else np.zeros_like(dimg)
    meta = dict(seed=int(seed), kind=case.kind, flux=case.flux, nvis=int(case.u.size), niter=info.niter,
                nfev=info.nfev, cost=info.cost, solve_s=info.wall_s, total_s=time.perf_counter() - t0,
                dirty_peak=peak, target_peak=tmax)
    return x, tgt, meta


def dihedral(a, k):
    """The k-th (0..7) symmetry of the square applied to the two leading
    image axes of a (transpose if k >= 4, then k % 4 quarter turns)."""
    if k >= 4:
        a = np.swapaxes(a, 0, 1)
    return np.rot90(a, k % 4, axes=(0, 1))
