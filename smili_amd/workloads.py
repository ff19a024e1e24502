"""Seeded imaging workloads at the sizes the benchmarks and tests use.

* ``single_large`` -- one large objective+gradient: 256x256 image at
  2 uas/pixel (fine grid 512x512, ES width 7), amplitudes on 1e5 uv points
  uniform in a disk of radius 10 Glambda, closure phases over all 220
  triangles and log closure amplitudes over all 495 quadrangles (2 each) of
  12 stations at 500 timestamps (230 GHz, dec +12), L1 + TV + TSV + KL on a
  seeded mix of 3-6 Gaussians.
* ``small_fcv`` -- 64x64 at 2 uas/pixel, a Gaussian-blurred ring (diameter
  40 uas, width 10 uas FWHM, 0.6 Jy), ~2,000 complex visibilities from 8
  stations over 12 h (100 scans), sigma = 1% of the zero-baseline flux;
  full-visibility chi-square + L1 (1) + TSV (1e4), non-negative, from a flat
  20 uas Gaussian.
"""
from __future__ import annotations

import math

import numpy as np

from . import synth
from .imaging.prep import prepare

UAS = synth.UAS


def single_large(seed=0, n=256, n_amp=100_000, n_st=12, n_times=500, cp_ca=True,
                 amp_radius=10e9, regs=("l1", "tv", "tsv", "kl")):
    rng = np.random.default_rng(seed)
    dx, dy = -2.0 * UAS, 2.0 * UAS
    sky = synth.random_gaussians(rng, 3, 6, total_flux=1.0, fov=n * 2.0 * UAS * 0.5)
    flux0 = sum(c.flux for c in sky)
    sigma = 0.01 * flux0
    ua, va = synth.disk_uv(rng, n_amp, amp_radius)
    amptab = synth.vis_table(rng, sky, ua, va, sigma)
    bstab = catab = None
    if cp_ca:
        scans = synth.earth_rotation_scans(rng, n_st, 12.0, 230e9, n_times, 12.0)
        bstab, catab = synth.closure_tables(rng, sky, scans, sigma)
    img = synth.sky_image(sky, n, n, dx, dy)
    lam = dict(l1_lambda=1.0 if "l1" in regs else -1, tv_lambda=1.0 if "tv" in regs else -1,
               tsv_lambda=1.0 if "tsv" in regs else -1, kl_lambda=1.0 if "kl" in regs else -1)
    prep = prepare(img, dx, dy, amptable=amptab, bstable=bstab, catable=catab, **lam)
    return prep, dict(sky=sky, image=img, dx=dx, dy=dy)


def small_fcv(seed=1, n=64, n_st=8, n_scans=100, elev_cut=10.0):
    rng = np.random.default_rng(seed)
    dx, dy = -2.0 * UAS, 2.0 * UAS
    sky = [synth.Ring(flux=0.6, diameter=40 * UAS, blur_fwhm=10 * UAS)]
    scans = synth.earth_rotation_scans(rng, n_st, 12.0, 230e9, n_scans, 12.0, elev_cut_deg=elev_cut,
                                       gst0=-6.0, lon_deg=(-70.0, 70.0), lat_deg=(-30.0, 60.0))
    u = np.concatenate([s.u for s in scans])
    v = np.concatenate([s.v for s in scans])
    vis = synth.vis_table(rng, sky, u, v, 0.01 * 0.6)
    init = synth.sky_image([synth.Gaussian(0.6, 20 * UAS)], n, n, dx, dy)
    prep = prepare(init, dx, dy, vistable=vis, l1_lambda=1.0, tsv_lambda=1e4)
    return prep, dict(sky=sky, image=synth.sky_image(sky, n, n, dx, dy), dx=dx, dy=dy, vis=vis)


def algorithmic_bytes(prep, Nx=None, Ny=None):
    """HBM bytes one objective+gradient evaluation must move at minimum:
    16 Npix (x in, grad out, FP64) + 32 G (fine grid complex64, two passes per
    transform, two transforms) + 16 M (uv' read by both transforms) + 16 M
    (model V write + read) + 32 M (FP64 residual write + read) + per-datum
    index/data/variance reads (fcv 28, amp 20, cp 28, ca 32 bytes)."""
    kw = prep.kwargs
    Nx = Nx or kw["Nx"]
    Ny = Ny or kw["Ny"]
    nf1 = 2 * Nx
    nf2 = 2 * Ny
    npix = np.asarray(kw["xidx"]).size
    M = np.asarray(kw["u"]).size
    nfcv = 0 if kw.get("uvidxfcv") is None else len(kw["uvidxfcv"])
    namp = 0 if kw.get("uvidxamp") is None else len(kw["uvidxamp"])
    ncp = 0 if kw.get("uvidxcp") is None else kw["uvidxcp"].shape[1]
    nca = 0 if kw.get("uvidxca") is None else kw["uvidxca"].shape[1]
    G = nf1 * nf2
    return int(16 * npix + 32 * G + 16 * M + 16 * M + 32 * M +
               28 * nfcv + 20 * namp + 28 * ncp + 32 * nca)
