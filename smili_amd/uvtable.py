"""Visibility-table operations around the imaging cost, with the semantics
of SMILI's VisTable methods (smili/uvdata/uvtable/vistable.py):

* ``comp / real / imag`` (323-340);
* ``eval_image / residual_image / chisq_image`` (361-433) through
  ``_call_fftlib`` (435-640): model_fcv / model_amp of an image or a movie;
* ``uvdensity / uvweight`` (2432-2537): uv-density (uniform) and error
  weights;
* ``fit_beam`` (2539-2625): Gaussian beam from the weighted second moments
  of the uv coverage;
* ``map_beam / map_residual / map_clean`` (2627-2800): dirty beam and
  residual map by a type-1 transform of the weighted visibilities;
* ``make_bstable / make_catable`` (743-975, 1532-1842) and the station
  helpers, from smili_amd.closures: closure-phase and log
  closure-amplitude tables built from a visibility table.

Tables are DataFrames or mappings of columns (u, v, amp, phase, sigma; and
frmidx for movies); images have ``header`` (nx, ny, dx, dy in degrees,
nxref, nyref) and ``data[istokes, ifreq, iy, ix]``; movies have ``Nt`` and
``images``.  Transforms run on the GPU (smili_amd.fftlib).
"""
from __future__ import annotations

import copy

import numpy as np
from scipy.spatial import cKDTree

from . import fftlib
from .closures import check_snr, make_bstable, make_catable, reorder_st_snr, station_dic, uvsort  # noqa: F401
from .imaging.prep import get_uvlist
from .imaging.prep3d import get_uvlist_loop


def _col(t, k):
    return np.asarray(t[k], np.float64)


def _copy_table(t):
    if hasattr(t, "copy") and hasattr(t, "columns"):
        return t.copy()
    return {k: np.array(v, copy=True) for k, v in t.items()}


def comp(table):
    return _col(table, "amp") * np.exp(1j * np.deg2rad(_col(table, "phase")))


def real(table):
    return _col(table, "amp") * np.cos(np.deg2rad(_col(table, "phase")))


def imag(table):
    return _col(table, "amp") * np.sin(np.deg2rad(_col(table, "phase")))


# ----------------------------------------------------------------------------
# model / residual tables (vistable.py:361-640)
# ----------------------------------------------------------------------------
def _call_fftlib(table, imfits, mask=None, amptable=False, istokes=0, ifreq=0, tune=None):
    """vistable.py:435-640 -> ((outputs of model_fcv or model_amp), Ndata),
    per-datum outputs in the table's row order."""
    is_movie = hasattr(imfits, "images")
    if is_movie:
        Nt = imfits.Nt
        image = imfits.images[0]
        Iin = [np.asarray(im.data, np.float64)[istokes, ifreq] for im in imfits.images]
    else:
        image = imfits
        Iin = [np.asarray(imfits.data, np.float64)[istokes, ifreq]]
    h = image.header
    Nx, Ny = int(h["nx"]), int(h["ny"])
    xg, yg = np.meshgrid(np.arange(Nx) + 1, np.arange(Ny) + 1)
    if mask is None:
        xidx, yidx = xg.reshape(-1), yg.reshape(-1)
        Iin = [I.reshape(-1) for I in Iin]
    else:
        idx = np.where(np.asarray(mask, bool))
        xidx, yidx = xg[idx], yg[idx]
        Iin = [I[idx] for I in Iin]
    dx_rad, dy_rad = np.deg2rad(h["dx"]), np.deg2rad(h["dy"])
    Nxref, Nyref = h["nxref"], h["nyref"]
    n = len(np.asarray(table["u"]))
    if is_movie:
        tabs = (None, table, None, None) if amptable else (table, None, None, None)
        u, v, ifcv, iamp, _, _, Nuvs, order = get_uvlist_loop(Nt, *tabs)
        rows = order[1] if amptable else order[0]
    else:
        tabs = (None, table, None, None) if amptable else (table, None, None, None)
        u, v, ifcv, iamp, _, _ = get_uvlist(*tabs)
        rows = np.arange(n)
    u = u * (2 * np.pi * dx_rad)
    v = v * (2 * np.pi * dy_rad)
    x = np.concatenate(Iin)
    geom = (x, xidx, yidx, Nxref, Nyref, Nx, Ny)
    if not amptable:
        ph = np.deg2rad(_col(table, "phase"))[rows]
        am = _col(table, "amp")[rows]
        var = np.square(_col(table, "sigma"))[rows]
        args = (ifcv, am * np.cos(ph), am * np.sin(ph), var)
        if is_movie:
            out = fftlib.fftlib3d.model_fcv(*geom, Nt, u, v, Nuvs, *args, tune=tune)
        else:
            out = fftlib.model_fcv(*geom, u, v, *args, tune=tune)
    else:
        args = (iamp, _col(table, "amp")[rows], np.square(_col(table, "sigma"))[rows])
        if is_movie:
            out = fftlib.fftlib3d.model_amp(*geom, Nt, u, v, Nuvs, *args, tune=tune)
        else:
            out = fftlib.model_amp(*geom, u, v, *args, tune=tune)
    # per-datum outputs back to the table's row order
    inv = np.empty(n, np.int64)
    inv[rows] = np.arange(n)
    out = tuple(o[inv] if (isinstance(o, np.ndarray) and o.shape == (n,) and k >= 2) else o
                for k, o in enumerate(out))
    return out, n


def eval_image(table, imfits, mask=None, amptable=False, istokes=0, ifreq=0, tune=None):
    """vistable.py:361-386: a copy of the table holding the model."""
    out, n = _call_fftlib(table, imfits, mask, amptable, istokes, ifreq, tune)
    t = _copy_table(table)
    if not amptable:
        fcv = out[2] + 1j * out[3]
        t["amp"] = np.abs(fcv)
        t["phase"] = np.angle(fcv, deg=True)
    else:
        t["amp"] = out[2]
        t["phase"] = np.zeros(n)
    return t


def residual_image(table, imfits, mask=None, amptable=False, istokes=0, ifreq=0, tune=None):
    """vistable.py:388-414: a copy of the table holding the residual."""
    out, n = _call_fftlib(table, imfits, mask, amptable, istokes, ifreq, tune)
    t = _copy_table(table)
    if not amptable:
        r = out[4] + 1j * out[5]
        t["amp"] = np.abs(r)
        t["phase"] = np.angle(r, deg=True)
    else:
        t["amp"] = out[3]
        t["phase"] = np.zeros(n)
    return t


def chisq_image(table, imfits, mask=None, amptable=False, istokes=0, ifreq=0, tune=None):
    """vistable.py:416-433 -> (chisq, reduced chisq)."""
    out, n = _call_fftlib(table, imfits, mask, amptable, istokes, ifreq, tune)
    chisq = out[0]
    return chisq, chisq / (2 * n if not amptable else n)


# ----------------------------------------------------------------------------
# weights and beams (vistable.py:2432-2625)
# ----------------------------------------------------------------------------
def uvdensity(table, image, npix=2):
    """vistable.py:2432-2477: for each point, the number of points strictly
    inside a circle of radius npix (in units of 1/fov, fov = twice the
    largest pixel radius) around it or around its conjugate."""
    x, y = _xygrid_rad(image)
    r = np.sqrt(x * x + y * y).max() * 2
    du = 1 / r
    u = _col(table, "u") / du
    v = _col(table, "v") / du
    pts = np.stack([u, v], axis=1)
    tree = cKDTree(pts)
    out = np.empty(u.size)
    for i in range(u.size):
        a = tree.query_ball_point([u[i], v[i]], npix)
        b = tree.query_ball_point([-u[i], -v[i]], npix)
        cand = np.unique(np.asarray(a + b, np.int64))
        du2p = (u[cand] - u[i]) ** 2 + (v[cand] - v[i]) ** 2
        du2m = (u[cand] + u[i]) ** 2 + (v[cand] + v[i]) ** 2
        out[i] = np.count_nonzero((du2p < npix ** 2) | (du2m < npix ** 2))
    return out


def _xygrid_rad(image):
    h = image.header
    xg = np.deg2rad(h["dx"]) * (np.arange(h["nx"]) - h["nxref"] + 1)
    yg = np.deg2rad(h["dy"]) * (np.arange(h["ny"]) - h["nyref"] + 1)
    return np.meshgrid(xg, yg)


def uvweight(table, uniform=True, image=None, werror=0, npix=2):
    """vistable.py:2479-2537: normalised product of density and error
    weights."""
    n = len(np.asarray(table["amp"]))
    if not uniform:
        uvw = np.ones(n)
    else:
        if image is None:
            raise ValueError("uniform weighting needs an image")
        uvw = 1.0 / uvdensity(table, image, npix)
        uvw /= uvw.max()
    if werror == 0:
        ew = np.ones(n)
    else:
        ew = np.power(_col(table, "sigma"), werror)
        ew /= ew.max()
    w = uvw * ew
    return w / w.sum()


def fit_beam(table, image, uniform=True, werror=0, npix=2, angunit=None, difmap_fudge=False):
    """vistable.py:2539-2625 -> dict(majsize, minsize, angunit, pa)."""
    from .imaging.image import _DEG_PER
    angunit = angunit or getattr(image, "angunit", "uas")
    w = uvweight(table, uniform, image, werror, npix)
    w = w / w.sum()
    u = _col(table, "u")
    v = _col(table, "v")
    muu = np.sum(u * u * w)
    mvv = np.sum(v * v * w)
    muv = np.sum(u * v * w)
    ftmp = np.sqrt(4.0 * muv * muv + (muu - mvv) ** 2)
    bpa = -0.5 * np.arctan2(2.0 * muv, muu - mvv) * 180. / np.pi
    varuv_maj = (muu + mvv) / 2. + ftmp / 2.
    varuv_min = (muu + mvv) / 2. - ftmp / 2.
    fudge = 0.7 / 2. if difmap_fudge else np.sqrt(8 * np.log(2)) / (2 * np.pi)
    conv = 180 / np.pi / _DEG_PER[angunit]
    return {"majsize": fudge / np.sqrt(varuv_min) * conv, "minsize": fudge / np.sqrt(varuv_maj) * conv,
            "angunit": angunit, "pa": bpa}


def _weighted_map(image, u, v, vis, w, istokes, ifreq):
    """Re A1 of the weighted visibilities on (u, -u) with the reference
    pixel's phase (vistable.py:2670-2694, 2740-2772), as an image copy."""
    h = image.header
    Nx, Ny = int(h["nx"]), int(h["ny"])
    dx_rad, dy_rad = np.deg2rad(h["dx"]), np.deg2rad(h["dy"])
    us = u * 2 * np.pi * dx_rad * -1
    vs = v * 2 * np.pi * dy_rad * -1
    ut = np.concatenate([us, -us])
    vt = np.concatenate([vs, -vs])
    wt = np.concatenate([w, w])
    wt = wt / wt.sum()
    V = np.concatenate([vis, np.conj(vis)])
    dix = Nx / 2. + 1 - h["nxref"]
    diy = Ny / 2. + 1 - h["nyref"]
    V = V * np.exp(1j * (ut * dix + vt * diy)) * wt
    m = fftlib.nufft_adj_real1d(ut, vt, V.real, V.imag, Nx, Ny).reshape(Ny, Nx)
    out = copy.deepcopy(image)
    data = np.array(out.data, dtype=np.float64, copy=True)
    data[istokes, ifreq] = m
    out.data = data
    if hasattr(out, "update_fits"):
        out.update_fits()
    return out


def map_beam(table, image, uniform=True, werror=0, npix=2, normalize=True, istokes=0, ifreq=0):
    """vistable.py:2627-2699: the synthesized (dirty) beam."""
    w = uvweight(table, uniform, image, werror, npix)
    n = w.size
    out = _weighted_map(image, _col(table, "u"), _col(table, "v"), np.ones(n, np.complex128), w,
                        istokes, ifreq)
    if normalize:
        d = np.asarray(out.data)
        d[istokes, ifreq] /= d[istokes, ifreq].max()
    return out


def map_residual(table, image, uniform=False, werror=-2, npix=2, istokes=0, ifreq=0, tune=None):
    """vistable.py:2701-2778: the residual (dirty) map of the data minus
    the image's model visibilities."""
    resid = residual_image(table, image, istokes=istokes, ifreq=ifreq, tune=tune)
    w = uvweight(table, uniform, image, werror, npix)
    return _weighted_map(image, _col(resid, "u"), _col(resid, "v"), comp(resid), w, istokes, ifreq)


def map_clean(table, image, restore=False, errorweight=-2, istokes=0, ifreq=0, tune=None):
    """vistable.py:2780-2801 without restoring (restore needs the image's
    beam convolution): residual map + model image.  ``errorweight`` is the
    error-weight index of map_residual (the reference passes it under a
    keyword map_residual does not take, vistable.py:2797)."""
    if restore:
        raise NotImplementedError("restore=True needs a beam convolution of the model image")
    res = map_residual(table, image, werror=errorweight, istokes=istokes, ifreq=ifreq, tune=tune)
    out = copy.deepcopy(image)
    data = np.array(out.data, dtype=np.float64, copy=True)
    data[istokes, ifreq] = np.asarray(res.data)[istokes, ifreq] + np.asarray(image.data)[istokes, ifreq]
    out.data = data
    if hasattr(out, "update_fits"):
        out.update_fits()
    return out


# ----------------------------------------------------------------------------
# CSV tables (vistable.py:3461-3481, bstable.py:1224-1244, catable.py:1196-1216)
# ----------------------------------------------------------------------------
def _read_table(filename, **args):
    import pandas as pd
    t = pd.read_csv(filename, **args)
    if "utc" in t.columns:
        t["utc"] = pd.to_datetime(t["utc"])
    return t


def read_vistable(filename, **args):
    """A visibility table from a CSV file (pd.read_csv; utc parsed to
    datetimes), as smili.uvdata.read_vistable (vistable.py:3461-3481).
    Columns the imager reads: u, v, amp, phase (deg), sigma (and frmidx for
    movies)."""
    return _read_table(filename, **args)


def read_bstable(filename, **args):
    """A closure-phase table from a CSV file (bstable.py:1224-1244); the
    imager reads u12, v12, u23, v23, u31, v31, amp, phase (deg), sigma."""
    return _read_table(filename, **args)


def read_catable(filename, **args):
    """A closure-amplitude table from a CSV file (catable.py:1196-1216); the
    imager reads u1..v4, logamp, logsigma."""
    return _read_table(filename, **args)
