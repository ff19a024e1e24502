"""Random small imaging problems in calc_cost's argument space, shared by the
oracle tests and the GPU parity tests."""
from __future__ import annotations

import dataclasses

import numpy as np

from oracle import smili_oracle as O


def random_problem(rng, Nx, Ny, M=300, nfcv=200, namp=150, ncp=120, nca=100,
                   regs=("l1", "tv", "tsv", "kl", "gs"), window=False, umax=np.pi * 0.95,
                   noise=0.05, extra_regs=()):
    """A problem whose data come from a random 'true' image plus noise, at
    random uv points (v >= 0, as the uv list holds), with random signed
    indices.  Returns (oracle.Problem, x)."""
    if window:
        mask = rng.uniform(size=(Ny, Nx)) < 0.7
        mask[Ny // 2, Nx // 2] = True
        yy, xx = np.nonzero(mask)
    else:
        yy, xx = np.mgrid[0:Ny, 0:Nx]
        yy = yy.reshape(-1)
        xx = xx.reshape(-1)
    xidx = (xx + 1).astype(np.int32)
    yidx = (yy + 1).astype(np.int32)
    npix = xidx.size
    u = rng.uniform(-umax, umax, M)
    v = rng.uniform(0, umax, M)
    truth = np.abs(rng.standard_normal(npix)) + 0.1
    I2d = O.image_1d_to_2d(truth, xidx, yidx, Nx, Ny)
    Vt = O.dft_fwd(I2d, u, v)
    scale = np.abs(Vt).mean()

    def sidx(n, k=None):
        shape = (n,) if k is None else (k, n)
        j = rng.integers(1, M + 1, size=shape)
        s = np.where(rng.uniform(size=shape) < 0.5, -1, 1)
        return (j * s).astype(np.int32)

    def model(idx):
        j = np.abs(idx) - 1
        return np.where(idx > 0, Vt[j], np.conj(Vt[j]))

    sd = noise if noise > 0 else 0.05          # variances stay positive without noise
    kw = dict(Nx=Nx, Ny=Ny, xidx=xidx, yidx=yidx, u=u, v=v,
              Nxref=Nx / 2.0 + 0.5, Nyref=Ny / 2.0 + 0.5)
    n_on = dict(fcv=nfcv, amp=namp, cp=ncp, ca=nca)
    wt = {k: (1.0 / max(2 * nfcv + namp + ncp + nca, 1)) for k in n_on}
    if nfcv:
        idx = sidx(nfcv)
        Vo = model(idx) + noise * scale * (rng.standard_normal(nfcv) + 1j * rng.standard_normal(nfcv))
        kw.update(uvidxfcv=idx, Vfcv=Vo, Varfcv=(sd * scale) ** 2 * rng.uniform(0.5, 2, nfcv), wfcv=wt["fcv"])
    if namp:
        idx = sidx(namp)
        Ao = np.abs(model(idx)) * (1 + noise * rng.standard_normal(namp))
        kw.update(uvidxamp=idx, Vamp=Ao, Varamp=(sd * scale) ** 2 * rng.uniform(0.5, 2, namp), wamp=wt["amp"])
    if ncp:
        idx = sidx(ncp, 3)
        ph = sum(np.angle(model(idx[k])) for k in range(3))
        kw.update(uvidxcp=idx, CP=ph + noise * rng.standard_normal(ncp),
                  Varcp=sd ** 2 * rng.uniform(0.5, 2, ncp), wcp=wt["cp"])
    if nca:
        idx = sidx(nca, 4)
        a = [np.log(np.abs(model(idx[k]))) for k in range(4)]
        kw.update(uvidxca=idx, CA=a[0] + a[1] - a[2] - a[3] + noise * rng.standard_normal(nca),
                  Varca=sd ** 2 * rng.uniform(0.5, 2, nca), wca=wt["ca"])
    for r in regs:
        lam = float(rng.uniform(0.1, 2.0))
        if r in ("kl", "gs"):
            kw[r + "_l"] = lam
            kw[r + "_wgt"] = rng.uniform(0.5, 1.5, npix)
        else:
            kw[r + "_l"] = lam
            kw[r + "_wgt"] = rng.uniform(0.5, 1.5, npix)
    for r in extra_regs:
        if r == "tfd":
            kw.update(tfd_l=0.3, tfd_tgtfd=float(truth.sum()))
        elif r == "cen":
            kw.update(cen_l=0.2, cen_alpha=2.0)
        elif r == "sm":
            kw.update(sm_l=0.01, sm_maj=6.0, sm_min=3.0, sm_phi=0.3)
    x = np.abs(truth * rng.uniform(0.7, 1.3, npix))
    prob = O.Problem(**kw)
    prob.extra["truth"] = truth
    return prob, x


def problem_kwargs(p: O.Problem):
    """oracle.Problem -> CostProblem keyword arguments (same names)."""
    d = {f.name: getattr(p, f.name) for f in dataclasses.fields(p) if f.name != "extra"}
    return {k: v for k, v in d.items() if v is not None}


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def random_movie(rng, Nz, Nx, Ny, nuvs, nfcv=300, namp=200, ncp=150, nca=120,
                 regs=("l1", "tv", "tsv", "kl", "gs"), window=False, umax=np.pi * 0.95,
                 noise=0.05, rt=True, ri=True, lc=False, lc_nidx=1, extra_regs=()):
    """A dynamic-imaging problem: Nz frames of a slowly changing random sky,
    frame z observed at nuvs[z] random uv points (0 allowed); the data index
    the concatenated uv list.  Returns (oracle.Movie, x[Nz*npix])."""
    nuvs = np.asarray(nuvs, np.int64)
    assert nuvs.size == Nz
    M = int(nuvs.sum())
    base, _ = random_problem(rng, Nx, Ny, M=max(M, 1), nfcv=0, namp=0, ncp=0, nca=0, regs=regs,
                             window=window, umax=umax, extra_regs=extra_regs)
    npix = base.npix
    truth0 = base.extra["truth"]
    drift = rng.standard_normal(npix) * 0.05
    truth = np.stack([np.abs(truth0 * (1 + 0.1 * np.sin(0.7 * z)) + drift * z) + 0.05 for z in range(Nz)])
    start = np.concatenate([[0], np.cumsum(nuvs)])
    Vt = np.zeros(M, np.complex128)
    for z in range(Nz):
        if nuvs[z]:
            s, e = start[z], start[z + 1]
            I2d = O.image_1d_to_2d(truth[z], base.xidx, base.yidx, Nx, Ny)
            Vt[s:e] = O.dft_fwd(I2d, base.u[s:e], base.v[s:e])
    live = np.nonzero(np.repeat(nuvs > 0, nuvs))[0] + 1   # all uv points belong to live frames
    scale = np.abs(Vt).mean()

    def sidx(n, k=None):
        shape = (n,) if k is None else (k, n)
        j = rng.choice(live, size=shape)
        s = np.where(rng.uniform(size=shape) < 0.5, -1, 1)
        return (j * s).astype(np.int32)

    def model(idx):
        j = np.abs(idx) - 1
        return np.where(idx > 0, Vt[j], np.conj(Vt[j]))

    kw = {}
    wt = 1.0 / (2 * nfcv + namp + ncp + nca)
    if nfcv:
        idx = sidx(nfcv)
        Vo = model(idx) + noise * scale * (rng.standard_normal(nfcv) + 1j * rng.standard_normal(nfcv))
        kw.update(uvidxfcv=idx, Vfcv=Vo, Varfcv=(noise * scale) ** 2 * rng.uniform(0.5, 2, nfcv), wfcv=wt)
    if namp:
        idx = sidx(namp)
        Ao = np.abs(model(idx)) * (1 + noise * rng.standard_normal(namp))
        kw.update(uvidxamp=idx, Vamp=Ao, Varamp=(noise * scale) ** 2 * rng.uniform(0.5, 2, namp), wamp=wt)
    if ncp:
        idx = sidx(ncp, 3)
        ph = sum(np.angle(model(idx[k])) for k in range(3))
        kw.update(uvidxcp=idx, CP=ph + noise * rng.standard_normal(ncp),
                  Varcp=noise ** 2 * rng.uniform(0.5, 2, ncp), wcp=wt)
    if nca:
        idx = sidx(nca, 4)
        a = [np.log(np.abs(model(idx[k]))) for k in range(4)]
        kw.update(uvidxca=idx, CA=a[0] + a[1] - a[2] - a[3] + noise * rng.standard_normal(nca),
                  Varca=noise ** 2 * rng.uniform(0.5, 2, nca), wca=wt)
    frame = dataclasses.replace(base, **kw)
    mv = O.Movie(frame=frame, Nuvs=nuvs)
    if rt:
        mv.rt_l = float(rng.uniform(0.5, 2.0))
        mv.rt_wgt = rng.uniform(0.5, 1.5, npix)
    if ri:
        mv.ri_l = float(rng.uniform(0.5, 2.0))
        mv.ri_wgt = rng.uniform(0.5, 1.5, npix)
    if lc:
        fl = truth.sum(axis=1)
        mv.lc_tgtfd = fl * rng.uniform(0.9, 1.1, Nz)
        mv.lc_l = rng.uniform(0.1, 0.5, Nz)
        mv.lc_nidx = lc_nidx
        frame.tfd_tgtfd = float(fl.mean())
    else:
        mv.lc_tgtfd = rng.uniform(0.8, 1.25, Nz)      # Rt / Ri normalisation still applies
    x = np.abs(truth * rng.uniform(0.8, 1.2, truth.shape)).reshape(-1)
    mv.extra["truth"] = truth.reshape(-1)
    return mv, x


def movie_kwargs(mv: O.Movie):
    """oracle.Movie -> MovieProblem keyword arguments."""
    kw = problem_kwargs(mv.frame)
    kw.update(Nz=mv.Nz, Nuvs=mv.Nuvs, rt_l=mv.rt_l, rt_wgt=mv.rt_wgt, ri_l=mv.ri_l, ri_wgt=mv.ri_wgt,
              lc_l=mv.lc_l, lc_tgtfd=mv.lc_tgtfd, lc_nidx=mv.lc_nidx)
    return kw
