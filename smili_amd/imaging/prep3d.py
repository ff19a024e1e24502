"""Problem set-up of SMILI's dynamic imager (movies), ahead of the GPU cost.

Mirrors what happens before the reference's 3-D hot loop:

* ``smili/imaging/lbfgs3d.py:256-660`` -- table sorting, total-flux guess,
  frame vectors, intensity normalisation over all frames, static
  regularizer weights, light curve, Rt / Ri weights, data arrays;
* ``smili/uvdata/uvtable/tools.py:255-336`` (get_uvlist_loop) -- one uv
  list per frame (tables carry a ``frmidx`` column), concatenated, with the
  uv indices offset by the points of the frames before;
* ``src/fortlib/fftim3d.f90:225-289`` -- data-weight normalisation and the
  phase shift of the full complex visibilities (as for a single image).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .prep import _data_kwargs, _static_regs, get_uvlist, init_tfdreg

_SORT_KEYS = ["utc", "stokesid", "ch", "st1", "st2"]


def _sorted(t):
    """lbfgs3d.py:262-269: DataFrames are sorted by utc, stokesid, ch, st1,
    st2 (when they carry those columns)."""
    if t is None or not hasattr(t, "sort_values"):
        return t
    if all(k in t.columns for k in _SORT_KEYS):
        return t.sort_values(by=_SORT_KEYS).reset_index(drop=True)
    return t


def _rows(t, idx):
    """Rows idx of a table (DataFrame or mapping of columns)."""
    if hasattr(t, "iloc"):
        return t.iloc[idx].reset_index(drop=True)
    return {k: np.asarray(v)[idx] for k, v in t.items()}


def _frames(t):
    return np.asarray(t["frmidx"]).astype(np.int64)


def get_uvlist_loop(Nt, fcvconcat=None, ampconcat=None, bsconcat=None, caconcat=None):
    """tools.py:255-336: per-frame uv lists (get_uvlist on the rows with
    frmidx == i) concatenated; indices shifted by the points of the earlier
    frames.  Returns (u, v, uvidxfcv, uvidxamp, uvidxcp, uvidxca, Nuvs,
    order) where ``order[k]`` lists the table rows in the order their
    indices are returned (frame by frame)."""
    tabs = (fcvconcat, ampconcat, bsconcat, caconcat)
    if all(t is None for t in tabs):
        raise ValueError("Error: No data are input.")
    us, vs = [], []
    idx = [[], [], [], []]
    order = [[], [], [], []]
    Nuvs = []
    idxcon = 0
    for i in range(Nt):
        single = []
        for k, t in enumerate(tabs):
            if t is None:
                single.append(None)
                continue
            r = np.nonzero(_frames(t) == i)[0]
            single.append(_rows(t, r) if r.size else None)
            if r.size:
                order[k].append(r)
        if all(s is None for s in single):
            Nuvs.append(0)
            continue
        u0, v0, i_fcv, i_amp, i_cp, i_ca = get_uvlist(*single)
        us.append(u0)
        vs.append(v0)
        Nuvs.append(u0.size)
        for k, a in enumerate((i_fcv, i_amp, i_cp, i_ca)):
            if single[k] is not None:
                idx[k].append(np.sign(a) * (np.abs(a) + idxcon))
        idxcon += u0.size
    u = np.concatenate(us)
    v = np.concatenate(vs)
    out = []
    for k, shape in enumerate(((1,), (1,), (3, 1), (4, 1))):
        if tabs[k] is None:
            out.append(np.zeros(shape, np.int32))
        elif k < 2:
            out.append(np.concatenate(idx[k]).astype(np.int32))
        else:
            out.append(np.hstack(idx[k]).astype(np.int32))
    order = [np.concatenate(o) if o else np.zeros(0, np.int64) for o in order]
    return (u, v, *out, np.asarray(Nuvs, np.int64), order)


@dataclass
class PreparedMovie:
    kwargs: dict              # MovieProblem arguments (fftim3d.calc_cost's argument list)
    x0: np.ndarray            # initial frames, frame-major, scaled by inormfactr
    inormfactr: float
    xidx: np.ndarray
    yidx: np.ndarray
    Nx: int
    Ny: int
    Nt: int


def _flat_or_prior(prior, winidx, Nyx, Npix):
    """Rt / Ri weight prior (lbfgs3d.py:566-578): flat 1/Npix, or the prior
    normalised to unit sum and smoothed as sqrt(p^2 + 1e-5^2)."""
    if prior is None:
        return np.full(Npix, 1.0 / Npix)
    a = np.asarray(prior, np.float64)
    p = a.reshape(Nyx).copy() if winidx is None else a[winidx].copy()
    p *= 1.0 / p.sum()
    return np.sqrt(p ** 2 + np.float64(1e-5) ** 2)


def prepare3d(frames, dx_rad, dy_rad, Nxref=None, Nyref=None, window=None,
              vistable=None, amptable=None, bstable=None, catable=None,
              w_vis=1, w_amp=1, w_bs=1, w_ca=1,
              l1_lambda=-1., l1_prior=None, l1_noise=1e-2, l1_type=1, l1_floor=1e-1,
              sm_lambda=-1, sm_maj=50., sm_min=50., sm_phi=0.,
              tv_lambda=-1., tv_prior=None, tsv_lambda=-1., tsv_prior=None,
              kl_lambda=-1., kl_prior=None, gs_lambda=-1, gs_prior=None,
              tfd_lambda=-1, tfd_tgterror=0.01,
              lc_lambda=-1, lc_array=None, lc_tgterror=0.01, lc_normalize=False,
              cen_lambda=-1, cen_alpha=3,
              rt_lambda=-1., rt_prior=None, ri_lambda=-1., ri_prior=None,
              totalflux=None, inorm=1.):
    """Arguments and defaults of smili.imaging.lbfgs3d.imaging3d
    (lbfgs3d.py:50-100) on plain arrays: ``frames`` [Nt, Ny, Nx] (Jy/pixel),
    ``dx_rad, dy_rad`` pixel increments in radians, ``window`` an optional
    boolean mask [Ny, Nx]; tables carry ``frmidx``.  sm_maj / sm_min are
    FWHMs in radians.

    Differences from lbfgs3d.py, where the reference fails:
    l1 without a prior uses the flat prior of the static imager (the
    reference leaves l1_wgt undefined there, lbfgs3d.py:356-359); tfd off
    without lc_array does not call len(None) (lbfgs3d.py:515)."""
    vistable, amptable, bstable, catable = map(_sorted, (vistable, amptable, bstable, catable))
    if vistable is None and amptable is None and bstable is None and catable is None:
        raise ValueError("Error: No data are input.")                      # lbfgs3d.py:256-260
    if vistable is None and amptable is None and totalflux is None and lc_array is None:
        raise ValueError("No absolute amplitude information: set totalflux / tfd_lambda")
    if vistable is None and amptable is None and (totalflux is None or tfd_lambda <= 0) and lc_array is None:
        raise ValueError("No absolute amplitude information: the total flux must be constrained")
    if totalflux is None and lc_array is not None:                        # lbfgs3d.py:283-291
        totalflux = float(np.median(lc_array))
    elif totalflux is None:
        totalflux = max(float(np.max(np.asarray(t["amp"], np.float64)))
                        for t in (vistable, amptable) if t is not None)
    F = np.asarray(frames, np.float64)
    Nt, Ny, Nx = F.shape
    Nyx = Nx * Ny
    Nxref = (Nx + 1) / 2.0 if Nxref is None else Nxref
    Nyref = (Ny + 1) / 2.0 if Nyref is None else Nyref
    xg, yg = np.meshgrid(np.arange(Nx) + 1, np.arange(Ny) + 1)
    if window is None:                                                      # lbfgs3d.py:311-330
        winidx = None
        Iin = [F[i].reshape(Nyx).copy() for i in range(Nt)]
        xidx = xg.reshape(Nyx)
        yidx = yg.reshape(Nyx)
    else:
        winidx = np.where(np.asarray(window, bool))
        Iin = [F[i][winidx].copy() for i in range(Nt)]
        xidx = xg[winidx]
        yidx = yg[winidx]
    Npix = Iin[0].size
    allI = np.concatenate(Iin)
    if inorm < 0:                                                           # lbfgs3d.py:337-346
        inormfactr = 1.0
    elif np.sum(allI) != 0:
        inormfactr = inorm / np.max(allI)
    else:
        inormfactr = inorm / (totalflux / Npix)
    tf = totalflux * inormfactr

    kw = dict(Nx=Nx, Ny=Ny, xidx=xidx, yidx=yidx, Nxref=Nxref, Nyref=Nyref)
    # static regularizers (lbfgs3d.py:349-503); sm_phi enters without the
    # sign flip of the static imager (lbfgs3d.py:399)
    kw.update(_static_regs(xidx, yidx, Nx, Ny, winidx, Nyx, tf, dx_rad,
                           l1_lambda, l1_prior, l1_noise, l1_type, l1_floor, tv_lambda, tv_prior,
                           tsv_lambda, tsv_prior, kl_lambda, kl_prior, gs_lambda, gs_prior,
                           cen_lambda, cen_alpha, sm_lambda, sm_maj, sm_min, sm_phi))
    lc_list = [] if lc_array is None else list(np.asarray(lc_array, np.float64).reshape(-1))
    if tfd_lambda <= 0:                                                     # lbfgs3d.py:512-524
        tfd_l, tfd_tgtfd = -1.0, 1.0
        if len(lc_list) > 0 and lc_normalize in ("static", "both"):
            tfd_tgtfd = float(np.mean(lc_list)) * inormfactr
    else:
        tfd_tgtfd = float(tf)
        tfd_l = init_tfdreg(tfd_lambda, tfd_tgtfd, tfd_tgterror)
    kw.update(tfd_l=tfd_l, tfd_tgtfd=tfd_tgtfd)
    # light curve (lbfgs3d.py:527-550)
    lc_nidx = -1
    if lc_lambda <= 0:
        lc_l = np.full(Nt, -1.0)
        lc_tgtfd = np.ones(Nt)
    else:
        if len(lc_list) != Nt:
            raise ValueError("Length of light array is inconsistent with time bin of the movie")
        lc_tgtfd = np.asarray(lc_list) * inormfactr
        lc_l = lc_lambda / (lc_tgterror * lc_tgtfd) ** 2
        lc_nidx = {"static": 1, "dynamic": 2, "both": 3}.get(lc_normalize, -1)
    # Rt / Ri (lbfgs3d.py:559-610)
    rt_wgt = ri_wgt = None
    if rt_lambda > 0:
        rt_wgt = 1.0 / (Nt * Npix * _flat_or_prior(rt_prior, winidx, Nyx, Npix))
    if ri_lambda > 0:
        ri_wgt = 1.0 / (Nt * Npix * _flat_or_prior(ri_prior, winidx, Nyx, Npix))
    kw.update(Nz=Nt, rt_l=float(rt_lambda), rt_wgt=rt_wgt, ri_l=float(ri_lambda), ri_wgt=ri_wgt,
              lc_l=lc_l, lc_tgtfd=lc_tgtfd, lc_nidx=lc_nidx)

    # uv list and data (lbfgs3d.py:662-738)
    if Nt > 1:
        u, v, ifcv, iamp, icp, ica, Nuvs, order = get_uvlist_loop(Nt, vistable, amptable, bstable, catable)
        # data rows in the order of their indices (frame by frame; for
        # frame-sorted tables this is the table order)
        tabs = [None if t is None else _rows(t, o)
                for t, o in zip((vistable, amptable, bstable, catable), order)]
    else:
        u, v, ifcv, iamp, icp, ica = get_uvlist(vistable, amptable, bstable, catable)
        Nuvs = np.array([u.size])
        tabs = [vistable, amptable, bstable, catable]
    kw.update(_data_kwargs(u, v, ifcv, iamp, icp, ica, *tabs, inormfactr, dx_rad, dy_rad,
                           w_vis, w_amp, w_bs, w_ca, Nx, Ny, Nxref, Nyref))
    kw["Nuvs"] = Nuvs
    return PreparedMovie(kw, allI * inormfactr, inormfactr, xidx, yidx, Nx, Ny, Nt)
