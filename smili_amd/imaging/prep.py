"""Problem set-up of SMILI's static imager, ahead of the GPU cost.

Mirrors what happens before the reference's hot loop:

* ``smili/imaging/lbfgs.py:202-489`` -- total-flux guess, image vector and
  pixel indices (x fastest), intensity normalisation factor, regularizer
  weights (``fortlib.image.init_*``), data arrays, uv list, uv scaling;
* ``src/fortlib/fftim2d.f90:178-241`` -- data-weight normalisation and the
  phase shift of the full complex visibilities to the image centre.

Tables are mappings of column name -> array (pandas DataFrames work), with
the reference's column names (``docs/examples/uvdata/*.csv``): vis/amp
tables ``u, v, amp, phase, sigma``; closure-phase tables
``u12, v12, u23, v23, u31, v31, amp, phase, sigma``; closure-amplitude
tables ``u1..v4, logamp, logsigma``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

ZEROEPS = 1e-10   # src/fortlib/image.f90:7


def _col(t, k):
    return np.asarray(t[k], dtype=np.float64)


# ----------------------------------------------------------------------------
# weight initialisers (src/fortlib/image.f90)
# ----------------------------------------------------------------------------
def smabs(x):
    return np.sqrt(np.square(x) + ZEROEPS)


def init_l1reg(prior):
    """image.f90:207-223: w = 1/smabs(p), normalised so sum w smabs(p) = 1."""
    w = 1.0 / smabs(prior)
    return w / np.sum(smabs(prior) * w)


def _tv_e(I2d):
    dx = np.zeros_like(I2d)
    dy = np.zeros_like(I2d)
    dx[:, :-1] = I2d[:, 1:] - I2d[:, :-1]
    dy[:-1, :] = I2d[1:, :] - I2d[:-1, :]
    return np.sqrt(dx * dx + dy * dy), dx * dx + dy * dy


def _to2d(I1d, xidx, yidx, Nx, Ny):
    I2d = np.zeros((Ny, Nx))
    I2d[yidx - 1, xidx - 1] = I1d
    return I2d


def init_tvreg(xidx, yidx, Nx, Ny, isflat, prior):
    """image.f90:409-446."""
    N = prior.size
    if isflat:
        return np.full(N, 1.0 / (ZEROEPS + prior.sum()) / N)
    e = _tv_e(_to2d(prior, xidx, yidx, Nx, Ny))[0][yidx - 1, xidx - 1]
    w = 1.0 / e
    return w / np.sum(e * w)


def init_tsvreg(xidx, yidx, Nx, Ny, isflat, prior):
    """image.f90:620-655."""
    N = prior.size
    if isflat:
        return np.full(N, 1.0 / (ZEROEPS + prior.sum() ** 2 / N) / N)
    w = 1.0 / (prior ** 2 + ZEROEPS)
    e = _tv_e(_to2d(prior, xidx, yidx, Nx, Ny))[1][yidx - 1, xidx - 1]
    return w / np.sum(e * w)


def init_klreg(kl_l_in, prior):
    """image.f90:278-291: (lambda / kl(P, P), P = smabs(prior))."""
    P = smabs(prior)
    a = smabs(P)                       # kl_e(I, P) smooths |I| (image.f90:325-331)
    kl_pp = np.sum(a * np.log(a / P) + P / math.e)
    return kl_l_in / kl_pp, P


def init_gsreg(prior):
    """image.f90:352-364."""
    return smabs(prior)


def init_tfdreg(tfd_l_in, tgtfd, tgter):
    """image.f90:760-771."""
    return tfd_l_in / (tgtfd * tgter) ** 2


# ----------------------------------------------------------------------------
# uv list (smili/uvdata/uvtable/tools.py:12-132)
# ----------------------------------------------------------------------------
def get_uvlist(fcvtable=None, amptable=None, bstable=None, catable=None):
    """Non-redundant uv list and signed 1-based indices per data type.
    Points with v < 0 are flipped to (-u, -v) with sign -1; the list is
    np.unique over u + iv."""
    if fcvtable is None and amptable is None and bstable is None and catable is None:
        raise ValueError("no data tables given")
    uvs, sgs = [], []

    def push(u, v):
        u = np.array(u, np.float64)
        v = np.array(v, np.float64)
        s = np.ones(u.size, np.int64)
        neg = v < 0
        u[neg] *= -1
        v[neg] *= -1
        s[neg] = -1
        uvs.append(u + 1j * v)
        sgs.append(s)

    n = [0, 0, 0, 0]
    if fcvtable is not None:
        push(_col(fcvtable, "u"), _col(fcvtable, "v"))
        n[0] = len(uvs[-1])
    if amptable is not None:
        push(_col(amptable, "u"), _col(amptable, "v"))
        n[1] = len(uvs[-1])
    if bstable is not None:
        for a, b in (("u12", "v12"), ("u23", "v23"), ("u31", "v31")):
            push(_col(bstable, a), _col(bstable, b))
        n[2] = len(uvs[-1])
    if catable is not None:
        for k in range(1, 5):
            push(_col(catable, "u%d" % k), _col(catable, "v%d" % k))
        n[3] = len(uvs[-1])
    uvstack = np.concatenate(uvs)
    sg = np.concatenate(sgs)
    uniq, inv = np.unique(uvstack, return_inverse=True)
    uvidx = (inv.reshape(-1) + 1) * sg
    o = 0
    fcv = uvidx[o:o + n[0]]
    o += n[0]
    amp = uvidx[o:o + n[1]]
    o += n[1]
    cp = uvidx[o:o + 3 * n[2]].reshape([n[2], 3], order="F").T
    o += 3 * n[2]
    ca = uvidx[o:o + 4 * n[3]].reshape([n[3], 4], order="F").T
    return (np.real(uniq).copy(), np.imag(uniq).copy(), fcv.astype(np.int32),
            amp.astype(np.int32), cp.astype(np.int32), ca.astype(np.int32))


# ----------------------------------------------------------------------------
# full set-up
# ----------------------------------------------------------------------------
@dataclass
class Prepared:
    kwargs: dict              # CostProblem arguments (calc_cost's argument list)
    x0: np.ndarray            # initial image vector, scaled by inormfactr
    inormfactr: float
    xidx: np.ndarray
    yidx: np.ndarray
    Nx: int
    Ny: int


def _prior_vec(prior, winidx, Nyx):
    if prior is None:
        return None
    a = np.asarray(prior, np.float64)
    return a.reshape(Nyx).copy() if winidx is None else a[winidx].copy()


def prepare(image, dx_rad, dy_rad, Nxref=None, Nyref=None, window=None,
            vistable=None, amptable=None, bstable=None, catable=None,
            w_vis=1, w_amp=1, w_bs=1, w_ca=1,
            l1_lambda=-1., l1_prior=None, l1_noise=1e-2,
            tv_lambda=-1, tv_prior=None, tsv_lambda=-1, tsv_prior=None,
            kl_lambda=-1, kl_prior=None, gs_lambda=-1, gs_prior=None,
            tfd_lambda=-1, tfd_tgterror=0.01, cen_lambda=-1, cen_alpha=3,
            sm_lambda=-1, sm_maj=0., sm_min=0., sm_phi=0.,
            totalflux=None, inorm=1.):
    """Arguments and defaults of smili.imaging.lbfgs.imaging (lbfgs.py:44-80)
    on plain arrays: ``image`` is the initial image [Ny, Nx] (Jy/pixel),
    ``dx_rad, dy_rad`` the pixel increments in radians (dx < 0 for RA),
    ``window`` an optional boolean mask [Ny, Nx] (imagewin)."""
    if vistable is None and amptable is None and bstable is None and catable is None:
        raise ValueError("Error: No data are input.")                     # lbfgs.py:186-190
    if vistable is None and amptable is None and (totalflux is None or tfd_lambda < 0):
        raise ValueError("No absolute amplitude information: set totalflux and tfd_lambda")
    if totalflux is None:                                                  # lbfgs.py:203-210
        totalflux = max(float(np.max(_col(t, "amp"))) for t in (vistable, amptable) if t is not None)
    Iimg = np.asarray(image, np.float64)
    Ny, Nx = Iimg.shape
    Nyx = Nx * Ny
    Nxref = (Nx + 1) / 2.0 if Nxref is None else Nxref
    Nyref = (Ny + 1) / 2.0 if Nyref is None else Nyref
    xg, yg = np.meshgrid(np.arange(Nx) + 1, np.arange(Ny) + 1)
    if window is None:                                                     # lbfgs.py:228-236
        winidx = None
        Iin = Iimg.reshape(Nyx).copy()
        xidx = xg.reshape(Nyx)
        yidx = yg.reshape(Nyx)
    else:                                                                  # lbfgs.py:237-249
        winidx = np.where(np.asarray(window, bool))
        Iin = Iimg[winidx].copy()
        xidx = xg[winidx]
        yidx = yg[winidx]
    Npix = Iin.size
    if inorm < 0:                                                          # lbfgs.py:252-260
        inormfactr = 1.0
    elif Iin.sum() != 0:
        inormfactr = inorm / Iin.max()
    else:
        inormfactr = inorm / (totalflux / Npix)
    tf = totalflux * inormfactr

    kw = dict(Nx=Nx, Ny=Ny, xidx=xidx, yidx=yidx, Nxref=Nxref, Nyref=Nyref)
    # regularizers (lbfgs.py:263-423)
    kw.update(_static_regs(xidx, yidx, Nx, Ny, winidx, Nyx, tf, dx_rad,
                           l1_lambda, l1_prior, l1_noise, 1, 0.0, tv_lambda, tv_prior, tsv_lambda, tsv_prior,
                           kl_lambda, kl_prior, gs_lambda, gs_prior, cen_lambda, cen_alpha,
                           sm_lambda, sm_maj, sm_min, -sm_phi))
    if tfd_lambda > 0:
        kw.update(tfd_l=init_tfdreg(tfd_lambda, tf, tfd_tgterror), tfd_tgtfd=tf)

    # data (lbfgs.py:429-489), weights and phase shift (fftim2d.f90:178-241)
    u, v, ifcv, iamp, icp, ica = get_uvlist(vistable, amptable, bstable, catable)
    kw.update(_data_kwargs(u, v, ifcv, iamp, icp, ica, vistable, amptable, bstable, catable,
                           inormfactr, dx_rad, dy_rad, w_vis, w_amp, w_bs, w_ca, Nx, Ny, Nxref, Nyref))
    return Prepared(kw, Iin * inormfactr, inormfactr, xidx, yidx, Nx, Ny)


def _static_regs(xidx, yidx, Nx, Ny, winidx, Nyx, tf, dx_rad,
                 l1_lambda, l1_prior, l1_noise, l1_type, l1_floor, tv_lambda, tv_prior, tsv_lambda, tsv_prior,
                 kl_lambda, kl_prior, gs_lambda, gs_prior, cen_lambda, cen_alpha,
                 sm_lambda, sm_maj, sm_min, sm_phi_deg):
    """Weights of the image-domain regularizers (lbfgs.py:263-423,
    lbfgs3d.py:349-512); ``tf`` is the scaled total flux, priors are
    [Ny, Nx] arrays or None (flat); sm_maj / sm_min are FWHMs in radians and
    sm_phi_deg the angle handed to the cost (sign as the caller's
    reference)."""
    Npix = xidx.size
    kw = {}
    if l1_lambda > 0:
        pr = _prior_vec(l1_prior, winidx, Nyx)
        if pr is None:
            pr = np.full(Npix, tf / Npix)
            pr *= tf / pr.sum()
        elif l1_type == 2:                                    # lbfgs3d.py:372-384
            pr *= tf / np.sum(pr)
            amax = np.abs(pr).max()
            pr[np.abs(pr) < amax * l1_noise] = amax * l1_floor
        else:
            amax = np.abs(pr).max()
            pr[np.abs(pr) < amax * l1_noise] = amax * l1_noise
            pr *= tf / pr.sum()
        kw.update(l1_l=l1_lambda, l1_wgt=init_l1reg(pr))
    if tv_lambda > 0:
        pr = _prior_vec(tv_prior, winidx, Nyx)
        flat = pr is None
        pr = np.full(Npix, tf / Npix) if flat else pr
        pr *= tf / pr.sum()
        kw.update(tv_l=tv_lambda, tv_wgt=init_tvreg(xidx, yidx, Nx, Ny, flat, pr))
    if tsv_lambda > 0:
        pr = _prior_vec(tsv_prior, winidx, Nyx)
        flat = pr is None
        pr = np.full(Npix, tf / Npix) if flat else pr
        pr *= tf / pr.sum()
        kw.update(tsv_l=tsv_lambda, tsv_wgt=init_tsvreg(xidx, yidx, Nx, Ny, flat, pr))
    if kl_lambda > 0:
        pr = _prior_vec(kl_prior, winidx, Nyx)
        pr = np.full(Npix, tf / Npix) if pr is None else pr
        pr *= tf / pr.sum()
        lam, w = init_klreg(kl_lambda, pr)
        kw.update(kl_l=lam, kl_wgt=w)
    if gs_lambda > 0:
        pr = _prior_vec(gs_prior, winidx, Nyx)
        pr = np.full(Npix, tf / Npix) if pr is None else pr
        pr *= tf / pr.sum()
        kw.update(gs_l=gs_lambda, gs_wgt=init_gsreg(pr))
    if cen_lambda > 0:
        kw.update(cen_l=cen_lambda, cen_alpha=cen_alpha)
    if sm_lambda > 0:
        dtheta = abs(dx_rad)
        kw.update(sm_l=sm_lambda,
                  sm_maj=(sm_maj / dtheta) ** 2 / (8. * math.log(2.)),
                  sm_min=(sm_min / dtheta) ** 2 / (8. * math.log(2.)),
                  sm_phi=sm_phi_deg / 180. * math.pi)
    return kw


def _data_kwargs(u, v, ifcv, iamp, icp, ica, vistable, amptable, bstable, catable,
                 inormfactr, dx_rad, dy_rad, w_vis, w_amp, w_bs, w_ca, Nx, Ny, Nxref, Nyref):
    """Data arrays (lbfgs.py:429-480), uv scaling (lbfgs.py:487-489), data
    weights (fftim2d.f90:178-213, fftim3d.f90:225-259) and the phase shift of
    the full complex visibilities to the image centre (fftim2d.f90:224-241,
    fftim3d.f90:270-289)."""
    kw = {}
    nf = na = nc = nq = 0
    if vistable is not None:
        ph = np.deg2rad(_col(vistable, "phase"))
        am = _col(vistable, "amp") * inormfactr
        Vfcv = am * np.cos(ph) + 1j * am * np.sin(ph)
        kw.update(uvidxfcv=ifcv, Vfcv=Vfcv, Varfcv=np.square(_col(vistable, "sigma")) * inormfactr ** 2)
        nf = ifcv.size
    if amptable is not None:
        kw.update(uvidxamp=iamp, Vamp=_col(amptable, "amp") * inormfactr,
                  Varamp=np.square(_col(amptable, "sigma")) * inormfactr ** 2)
        na = iamp.size
    if bstable is not None:
        kw.update(uvidxcp=icp, CP=np.deg2rad(_col(bstable, "phase")),
                  Varcp=np.square(_col(bstable, "sigma") / _col(bstable, "amp")))
        nc = icp.shape[1]
    if catable is not None:
        kw.update(uvidxca=ica, CA=_col(catable, "logamp"), Varca=np.square(_col(catable, "logsigma")))
        nq = ica.shape[1]
    u = u * (2 * np.pi * dx_rad)
    v = v * (2 * np.pi * dy_rad)
    kw.update(u=u, v=v)
    wtotal = 2 * nf * w_vis + na * w_amp + nc * w_bs + nq * w_ca
    kw.update(wfcv=w_vis / wtotal if nf else 0.0, wamp=w_amp / wtotal if na else 0.0,
              wcp=w_bs / wtotal if nc else 0.0, wca=w_ca / wtotal if nq else 0.0)
    if nf:
        s = np.sign(ifcv)
        j = np.abs(ifcv) - 1
        dix = Nxref - Nx / 2.0 - 1.0
        diy = Nyref - Ny / 2.0 - 1.0
        kw["Vfcv"] = kw["Vfcv"] * np.exp(1j * (u[j] * s * dix + v[j] * s * diy))
    return kw
