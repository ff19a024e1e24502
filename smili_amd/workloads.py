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
* ``dynamic_movie`` -- 64 frames of 128x128 at 2 uas/pixel; 3e5 complex
  visibilities in total, 2e3..8e3 per frame, from 10 seeded stations over
  12 h; a ring (diameter 40 uas) whose brightness asymmetry angle rotates by
  90 degrees across the frames; per-frame visibility chi-square + L1 + TSV +
  Rt, one joint problem.
* ``survey`` -- 512 independent problems: image sizes uniform over
  {64, 96, 128, 192, 256}^2, visibility counts log-uniform in [2e3, 5e4]
  from 6..12 seeded stations (Earth rotation, 12 h), L1 log-spaced over
  [1e-2, 1e2] and TSV over [1e2, 1e6] (8 x 8), TV and MEM (KL) each on/off,
  two seeded skies per setting.
"""
from __future__ import annotations

import math

import numpy as np

from . import synth
from .imaging.prep import prepare
from .imaging.prep3d import prepare3d

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


def _frame_counts(rng, n_frames, total, lo, hi):
    """Per-frame counts in [lo, hi] summing to total (uniform draws scaled,
    clipped, and the remainder spread over frames with room)."""
    c = rng.uniform(lo, hi, n_frames)
    c = np.clip(np.round(c * total / c.sum()), lo, hi).astype(np.int64)
    while c.sum() != total:
        d = total - int(c.sum())
        room = np.nonzero(c < hi)[0] if d > 0 else np.nonzero(c > lo)[0]
        k = room[: min(abs(d), room.size)]
        c[k] += 1 if d > 0 else -1
    return c


def dynamic_movie(seed=5, n_frames=64, n=128, n_vis=300_000, n_st=10, hours=12.0,
                  per_frame=(2000, 8000), rt_lambda=1.0, l1_lambda=1.0, tsv_lambda=1e4):
    rng = np.random.default_rng(seed)
    dx, dy = -2.0 * UAS, 2.0 * UAS
    nbl = n_st * (n_st - 1) // 2
    scans_per_frame = -(-per_frame[1] // nbl) + 1
    scans = synth.earth_rotation_scans(rng, n_st, 12.0, 230e9, n_frames * scans_per_frame, hours)
    counts = _frame_counts(rng, n_frames, n_vis, *per_frame)
    phi0 = float(rng.uniform(0, 2 * math.pi))
    skies, cols = [], {k: [] for k in ("u", "v", "amp", "phase", "sigma", "frmidx", "utc")}
    flux = 0.6
    for z in range(n_frames):
        sky = synth.asymmetric_ring(flux, 40 * UAS, 10 * UAS, phi0 + 0.5 * math.pi * z / max(n_frames - 1, 1))
        skies.append(sky)
        fs = scans[z * scans_per_frame:(z + 1) * scans_per_frame]
        u = np.concatenate([s.u for s in fs])
        v = np.concatenate([s.v for s in fs])
        t = np.concatenate([np.full(s.u.size, s.gst) for s in fs])
        pick = np.sort(rng.choice(u.size, int(counts[z]), replace=False))
        tab = synth.vis_table(rng, sky, u[pick], v[pick], 0.01 * flux)
        for k in ("u", "v", "amp", "phase", "sigma"):
            cols[k].append(tab[k])
        cols["frmidx"].append(np.full(pick.size, z))
        cols["utc"].append(t[pick])
    vis = {k: np.concatenate(a) for k, a in cols.items()}
    init = synth.sky_image([synth.Gaussian(flux, 20 * UAS)], n, n, dx, dy)
    frames0 = np.repeat(init[None], n_frames, axis=0)
    prep = prepare3d(frames0, dx, dy, vistable=vis, l1_lambda=l1_lambda, tsv_lambda=tsv_lambda,
                     rt_lambda=rt_lambda)
    truth = np.stack([synth.sky_image(s, n, n, dx, dy) for s in skies])
    return prep, dict(skies=skies, frames=truth, dx=dx, dy=dy, vis=vis, counts=counts)


SURVEY_SIZES = (64, 96, 128, 192, 256)


def survey(seed=3, n_problems=512, n_vis=(2_000, 50_000), stations=(6, 12)):
    """Returns (list of Prepared, meta).  Settings: the 8 x 8 grid of
    (L1, TSV) x TV on/off x KL on/off = 256, each with two skies."""
    rng = np.random.default_rng(seed)
    l1s = np.logspace(-2, 2, 8)
    tsvs = np.logspace(2, 6, 8)
    settings = [(a, b, tv, kl) for a in l1s for b in tsvs for tv in (False, True) for kl in (False, True)]
    preps, meta = [], []
    for i in range(n_problems):
        l1, tsv, tv, kl = settings[i % len(settings)]
        n = int(rng.choice(SURVEY_SIZES))
        nv = int(round(math.exp(rng.uniform(math.log(n_vis[0]), math.log(n_vis[1])))))
        n_st = int(rng.integers(stations[0], stations[1] + 1))
        nbl = n_st * (n_st - 1) // 2
        scans = synth.earth_rotation_scans(rng, n_st, float(rng.uniform(-30, 60)), 230e9,
                                           -(-nv // nbl) + 1, 12.0)
        u = np.concatenate([s.u for s in scans])
        v = np.concatenate([s.v for s in scans])
        pick = np.sort(rng.choice(u.size, nv, replace=False))
        dx, dy = -2.0 * UAS, 2.0 * UAS
        sky = synth.random_gaussians(rng, 3, 6, total_flux=float(rng.uniform(0.3, 2.0)),
                                     fov=min(n * 2.0 * UAS * 0.5, 150 * UAS))
        flux = sum(c.flux for c in sky)
        vis = synth.vis_table(rng, sky, u[pick], v[pick], 0.01 * flux)
        init = synth.sky_image([synth.Gaussian(flux, 30 * UAS)], n, n, dx, dy)
        preps.append(prepare(init, dx, dy, vistable=vis, l1_lambda=float(l1), tsv_lambda=float(tsv),
                             tv_lambda=1.0 if tv else -1, kl_lambda=1.0 if kl else -1))
        meta.append(dict(n=n, nvis=nv, stations=n_st, l1=float(l1), tsv=float(tsv), tv=tv, kl=kl))
    return preps, meta
