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
    wt = {k: (1.0 / (2 * nfcv + namp + ncp + nca)) for k in n_on}
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
