"""Pulse-function support of the static imager (``imaging(pulse_fwhm=...)``,
smili/imaging/lbfgs.py:436-478, 574-575, 1114-1194).

With a circular Gaussian pulse of FWHM theta the image is modelled as the
sky convolved with the pulse, so the tables are divided by the pulse's
Fourier transform before the solve,

    K(rho) = exp(-pi^2 theta^2 rho^2 / (4 ln 2)),   rho = the baseline length
                                                   (uvdist, in wavelengths),

and the solution is convolved with the pulse afterwards (total flux kept).
Visibility / amplitude tables: amp and sigma / K.  Closure phases: amp and
sigma / (K12 K23 K31), which leaves sigma / amp -- the closure-phase
variance the cost uses -- unchanged.  Closure amplitudes: amp and sigma /
(K1 K2 / (K3 K4)) and logamp - log(K1 K2 / (K3 K4)) (logsigma unchanged).
Tables are DataFrames or dicts of columns; copies are returned.  Baseline
lengths come from the uvdist* columns, or from u, v where a table lacks
them.
"""
from __future__ import annotations

import copy
import math

import numpy as np


def pulse_ft(fwhm_rad, uvdist):
    """Fourier transform of a unit-area circular Gaussian of FWHM fwhm_rad at
    baseline length uvdist (wavelengths)."""
    return np.exp(-(math.pi ** 2) * (fwhm_rad * np.asarray(uvdist, np.float64)) ** 2 / (4.0 * math.log(2.0)))


def _col(t, k):
    return np.asarray(t[k], dtype=np.float64)


def _uvdist(t, suffix=""):
    try:
        return _col(t, "uvdist" + suffix)
    except (KeyError, IndexError, ValueError):
        return np.hypot(_col(t, "u" + suffix), _col(t, "v" + suffix))


def _copy(t):
    return t.copy(deep=True) if hasattr(t, "copy") and hasattr(t, "columns") else copy.deepcopy(t)


def _set(t, k, val):
    t[k] = val if hasattr(t, "columns") else np.asarray(val)


def deburr_vistable(table, fwhm_rad):
    """lbfgs.py:1114-1135: amp, sigma / K(uvdist)."""
    out = _copy(table)
    k = pulse_ft(fwhm_rad, _uvdist(table))
    _set(out, "amp", _col(table, "amp") / k)
    _set(out, "sigma", _col(table, "sigma") / k)
    return out


def deburr_bstable(table, fwhm_rad):
    """lbfgs.py:1138-1165: amp, sigma / (K12 K23 K31)."""
    out = _copy(table)
    k = pulse_ft(fwhm_rad, _uvdist(table, "12")) * pulse_ft(fwhm_rad, _uvdist(table, "23")) * \
        pulse_ft(fwhm_rad, _uvdist(table, "31"))
    _set(out, "amp", _col(table, "amp") / k)
    _set(out, "sigma", _col(table, "sigma") / k)
    return out


def deburr_catable(table, fwhm_rad):
    """lbfgs.py:1168-1194: amp, sigma / (K1 K2 / (K3 K4)); logamp -
    log(K1 K2 / (K3 K4))."""
    out = _copy(table)
    k = [pulse_ft(fwhm_rad, _uvdist(table, str(n))) for n in range(1, 5)]
    r = k[0] * k[1] / k[2] / k[3]
    _set(out, "amp", _col(table, "amp") / r)
    _set(out, "logamp", _col(table, "logamp") - np.log(r))
    _set(out, "sigma", _col(table, "sigma") / r)
    return out


def convolve_gauss(data2d, fwhm_pix_x, fwhm_pix_y=None, save_totalflux=True):
    """Linear (zero-filled) convolution of an [Ny, Nx] image with a circular
    Gaussian of FWHM fwhm_pix (pixels), the kernel sampled on the image's own
    grid at offsets (i - (n - 1)/2) pixels and normalised to unit sum, its
    element n // 2 taken as the origin (the conventions of the convolution
    the reference's IMFITS.convolve_gauss calls, imdata.py:1705-1774); with
    save_totalflux the result is rescaled to the input's total flux."""
    img = np.asarray(data2d, dtype=np.float64)
    ny, nx = img.shape
    fy = fwhm_pix_x if fwhm_pix_y is None else fwhm_pix_y
    X = np.arange(nx) - (nx - 1) / 2.0
    Y = np.arange(ny) - (ny - 1) / 2.0
    sx = fwhm_pix_x / math.sqrt(2.0 * math.log(2.0)) / 2.0
    sy = fy / math.sqrt(2.0 * math.log(2.0)) / 2.0
    kern = np.exp(-X[None, :] ** 2 / (2 * sx * sx) - Y[:, None] ** 2 / (2 * sy * sy))
    kern /= kern.sum()
    # zero-filled linear convolution on a padded grid; kernel element
    # (ny // 2, nx // 2) at the origin
    py, px = 2 * ny, 2 * nx
    a = np.zeros((py, px))
    a[:ny, :nx] = img
    k = np.zeros((py, px))
    k[:ny, :nx] = kern
    k = np.roll(k, (-(ny // 2), -(nx // 2)), axis=(0, 1))
    out = np.fft.irfft2(np.fft.rfft2(a) * np.fft.rfft2(k), s=(py, px))[:ny, :nx]
    if save_totalflux:
        s = out.sum()
        if s != 0:
            out *= img.sum() / s
    return out
