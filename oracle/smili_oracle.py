"""FP64 CPU oracle for SMILI's regularized maximum-likelihood imaging cost.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``smili_amd``) never routes through it.

It restates, in plain numpy FP64, the reference algorithm of
``fftim2d.calc_cost`` (src/fortlib/fftim2d.f90:387-607) and its callees:

* forward model      -- exact direct DFT with FINUFFT's type-2 conventions
                        (src/wrapfinufft/wrapfinufft.f90:94-128: iflag=0 => +i,
                        CMCL mode order fk(-ms/2:(ms-1)/2), real image promoted)
* adjoint            -- exact direct DFT with FINUFFT's type-1 conventions
                        (wrapfinufft.f90:172-206: iflag=1 => +i) and the
                        Re(A Vresre) + Im(A Vresim) combination of
                        src/fortlib/fftlib.f90:152-174
* chi-square terms   -- src/fortlib/fftlib.f90:384-655
* regularizers       -- src/fortlib/image.f90:12-200, 207-754, 760-1158
* problem set-up     -- src/fortlib/fftim2d.f90:178-241 and
                        smili/imaging/lbfgs.py:202-489
* uv list            -- smili/uvdata/uvtable/tools.py:12-132
* dynamic imaging    -- fftim3d.calc_cost (src/fortlib/fftim3d.f90:472-640):
                        calc_chisq3d (src/fortlib/fftlib3d.f90:14-201),
                        calc_cost_reg3d and calc_cost_dynamical
                        (src/fortlib/image3d.f90:15-440)

A second forward/adjoint pair, ``es_nufft_type2`` / ``es_nufft_type1``,
implements the exponential-of-semicircle (ES) spreading NUFFT of Barnett,
Magland & af Klinteberg (SIAM J. Sci. Comput. 41(5), 2019), the published
method inside FINUFFT (an unpinned, un-vendored dependency of the reference,
src/wrapfinufft/wrapfinufft.f90:11-41).  It is pinned against the direct DFT
above, and it is what the CPU baseline times.

Parity pins (see tests/test_oracle.py): the DFT is checked against brute-force
loops, against the reference's own DFT sign convention (dftlib.f90:44-50),
against the adjoint identity and against closed-form Gaussian visibilities;
each chi-square gradient against central finite differences and against zero
residual at noise-free data; the regularizers against brute-force per-pixel
loops written from image.f90 and against finite differences; the TV reset
quirk (image.f90:577-578, 603-604) against a hand-built flat patch.  The
3-D cost is pinned by reductions to the 2-D one (one frame; identical frames
sharing one uv list), by finite differences of Rt / Ri (and the missing Rt
gradient of the last frame, image3d.f90:374), by closed forms on constant
and linear-in-time movies, and by the light-curve weight rescaling.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

# image.f90:7
ZEROEPS = 1e-10


# ----------------------------------------------------------------------------
# smoothed absolute value  (image.f90:1146-1158)
# ----------------------------------------------------------------------------
def smabs(x):
    return np.sqrt(np.square(x) + ZEROEPS)


def smabs_diff(x):
    return x / smabs(x)


# ----------------------------------------------------------------------------
# Direct DFTs with the reference's NUFFT conventions
# ----------------------------------------------------------------------------
def mode_indices(n):
    """CMCL mode order of FINUFFT (wrapfinufft.f90:19,31): pixel ix (1-based)
    carries mode k = ix - 1 - n//2, i.e. k in [-n/2, (n-1)/2]."""
    return np.arange(n, dtype=np.float64) - (n // 2)


def dft_fwd(I2d, u, v, chunk=4096):
    """Type-2 direct DFT  V_j = sum_{iy,ix} I[iy,ix] exp(+i(kx u_j + ky v_j)).

    I2d is numpy [Ny, Nx] (x fastest, = Fortran I2d(Nx,Ny)); u, v are the
    scaled coordinates u' = 2 pi dx_rad u (lbfgs.py:487-489).
    Follows FINUFFT_fwd, wrapfinufft.f90:94-128 (iflag=0 -> +i).
    """
    I2d = np.asarray(I2d, dtype=np.float64)
    Ny, Nx = I2d.shape
    kx = mode_indices(Nx)
    ky = mode_indices(Ny)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    out = np.empty(u.size, dtype=np.complex128)
    for s in range(0, u.size, chunk):
        e = slice(s, s + chunk)
        Ex = np.exp(1j * np.outer(u[e], kx))     # [m, Nx]
        Ey = np.exp(1j * np.outer(v[e], ky))     # [m, Ny]
        T = Ex @ I2d.T                           # [m, Ny]
        out[e] = np.sum(Ey * T, axis=1)
    return out


def dft_adj(c, u, v, Nx, Ny, chunk=4096):
    """Type-1 direct DFT  F[iy,ix] = sum_j c_j exp(+i(kx u_j + ky v_j)).

    Follows FINUFFT_adj, wrapfinufft.f90:172-206 (iflag=1 -> +i).
    Returns complex [Ny, Nx].
    """
    kx = mode_indices(Nx)
    ky = mode_indices(Ny)
    c = np.asarray(c, dtype=np.complex128)
    out = np.zeros((Ny, Nx), dtype=np.complex128)
    for s in range(0, c.size, chunk):
        e = slice(s, s + chunk)
        Ex = np.exp(1j * np.outer(u[e], kx))
        Ey = np.exp(1j * np.outer(v[e], ky))
        out += (Ey * c[e, None]).T @ Ex
    return out


def nufft_adj_resid(Vresre, Vresim, u, v, Nx, Ny, adj=None):
    """grad2d = Re(A1 Vresre) + Im(A1 Vresim)   (fftlib.f90:152-174)."""
    adj = adj or dft_adj
    return (adj(Vresre, u, v, Nx, Ny).real + adj(Vresim, u, v, Nx, Ny).imag)


# ----------------------------------------------------------------------------
# ES-kernel NUFFT (published method; what FINUFFT implements)
# ----------------------------------------------------------------------------
ES_WIDTH = 7            # ns = ceil(-log10(eps/10)) at eps = 1e-6
ES_BETA = 2.30 * ES_WIDTH  # beta/ns = 2.30 at upsampling factor 2
UPSAMP = 2


def es_kernel(z, beta=ES_BETA):
    """phi(z) = exp(beta (sqrt(1 - z^2) - 1)) on |z| <= 1, written in the
    cancellation-free form  -beta z^2 / (1 + sqrt(1 - z^2))."""
    z = np.asarray(z, dtype=np.float64)
    inside = np.abs(z) <= 1.0
    zz = np.where(inside, z, 0.0)
    s = np.sqrt(1.0 - zz * zz)
    val = np.exp(-beta * zz * zz / (1.0 + s))
    return np.where(inside, val, 0.0)


def es_correction(n, nf, w=ES_WIDTH, beta=ES_BETA, nquad=200):
    """Per-mode deconvolution factor c(k) = 1 / ((w/2) int_{-1}^{1} phi(z)
    cos(pi k w z / nf) dz) for k = -n/2 .. (n-1)/2 (Gauss-Legendre)."""
    x, wq = np.polynomial.legendre.leggauss(nquad)
    k = mode_indices(n)
    phi = es_kernel(x, beta)
    ft = (w / 2.0) * np.sum(wq[None, :] * phi[None, :] *
                            np.cos(np.pi * k[:, None] * w * x[None, :] / nf), axis=1)
    return 1.0 / ft


def next_smooth(n):
    """Smallest 2,3,5-smooth integer >= n."""
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def fine_grid_size(n, w=ES_WIDTH):
    return next_smooth(max(UPSAMP * n, 2 * w))


def _es_taps(x, nf, w=ES_WIDTH, beta=ES_BETA):
    """Grid start index l0 (mod nf) and the w tap weights for points x (rad)."""
    h = 2.0 * np.pi / nf
    t = np.mod(x / h, nf)                      # grid units in [0, nf)
    l0 = np.ceil(t - w / 2.0)
    d = t - l0                                 # distance to first tap, in (w/2-1, w/2]
    taps = np.arange(w, dtype=np.float64)
    z = (d[:, None] - taps[None, :]) * (2.0 / w)
    wts = es_kernel(z, beta)
    idx = np.mod(l0.astype(np.int64)[:, None] + np.arange(w)[None, :], nf)
    return idx, wts


def es_nufft_type2(I2d, u, v, w=ES_WIDTH, beta=ES_BETA):
    """ES-kernel type-2 NUFFT in FP64 (approximates dft_fwd to ~1e-7)."""
    I2d = np.asarray(I2d, dtype=np.float64)
    Ny, Nx = I2d.shape
    nf1, nf2 = fine_grid_size(Nx, w), fine_grid_size(Ny, w)
    cx = es_correction(Nx, nf1, w, beta)
    cy = es_correction(Ny, nf2, w, beta)
    g = np.zeros((nf2, nf1), dtype=np.complex128)
    ix = np.mod(mode_indices(Nx).astype(np.int64), nf1)
    iy = np.mod(mode_indices(Ny).astype(np.int64), nf2)
    g[np.ix_(iy, ix)] = I2d * cy[:, None] * cx[None, :]
    b = np.fft.ifft2(g) * (nf1 * nf2)          # sum_k g_k exp(+i k l h)
    xi, xw = _es_taps(np.asarray(u, np.float64), nf1, w, beta)
    yi, yw = _es_taps(np.asarray(v, np.float64), nf2, w, beta)
    bf = b.ravel()
    out = np.zeros(xi.shape[0], dtype=np.complex128)
    for a in range(w):
        base = yi[:, a] * nf1
        acc = np.zeros(xi.shape[0], dtype=np.complex128)
        for c in range(w):
            acc += bf[base + xi[:, c]] * xw[:, c]
        out += acc * yw[:, a]
    return out


def es_nufft_type1(c, u, v, Nx, Ny, w=ES_WIDTH, beta=ES_BETA):
    """ES-kernel type-1 NUFFT in FP64 (approximates dft_adj)."""
    nf1, nf2 = fine_grid_size(Nx, w), fine_grid_size(Ny, w)
    cx = es_correction(Nx, nf1, w, beta)
    cy = es_correction(Ny, nf2, w, beta)
    c = np.asarray(c, dtype=np.complex128)
    xi, xw = _es_taps(np.asarray(u, np.float64), nf1, w, beta)
    yi, yw = _es_taps(np.asarray(v, np.float64), nf2, w, beta)
    idx = (yi[:, :, None] * nf1 + xi[:, None, :]).reshape(-1)
    wt = (yw[:, :, None] * xw[:, None, :]).reshape(c.size, -1)
    g = (np.bincount(idx, weights=(wt * c.real[:, None]).reshape(-1), minlength=nf1 * nf2) +
         1j * np.bincount(idx, weights=(wt * c.imag[:, None]).reshape(-1), minlength=nf1 * nf2))
    g = g.reshape(nf2, nf1)
    F = np.fft.ifft2(g) * (nf1 * nf2)
    ix = np.mod(mode_indices(Nx).astype(np.int64), nf1)
    iy = np.mod(mode_indices(Ny).astype(np.int64), nf2)
    return F[np.ix_(iy, ix)] * cy[:, None] * cx[None, :]


def _scatter_add(arr, idx, vals):
    """arr[idx] += vals with repeated indices accumulated (the OpenMP array
    REDUCTION(+:Vresre,Vresim) of the reference loops)."""
    arr += np.bincount(np.asarray(idx).reshape(-1), weights=np.asarray(vals, np.float64).reshape(-1),
                       minlength=arr.size)


# ----------------------------------------------------------------------------
# Chi-square terms and residual vectors  (fftlib.f90:384-655)
# Vresre / Vresim are real arrays: the reference declares them complex but
# only ever adds real values into them (fftlib.f90:334-335, 433-434, ...).
# ----------------------------------------------------------------------------
def chisq_fcv(V, uvidx, Vobs, var, fnorm, Vresre, Vresim):
    """fftlib.f90:384-439."""
    j = np.abs(uvidx) - 1
    s = np.sign(uvidx)
    Vm = np.where(s > 0, V[j], np.conj(V[j]))
    r = Vobs - Vm
    chisq = np.sum(np.abs(r) ** 2 / var * fnorm)
    factor = -2.0 / var * fnorm
    _scatter_add(Vresre, j, factor * r.real)
    _scatter_add(Vresim, j, factor * r.imag * s)
    return chisq


def chisq_amp(V, uvidx, amp, var, fnorm, Vresre, Vresim):
    """fftlib.f90:442-493."""
    j = np.abs(uvidx) - 1
    model = np.abs(V[j])
    r = amp - model
    chisq = np.sum(r ** 2 / var * fnorm)
    factor = -2.0 * r / var / model * fnorm
    _scatter_add(Vresre, j, factor * V[j].real)
    _scatter_add(Vresim, j, factor * V[j].imag)
    return chisq


def chisq_ca(V, uvidx4, ca, var, fnorm, Vresre, Vresim):
    """fftlib.f90:496-573.  uvidx4 is [4, Nca] (Fortran uvidxca(4,Nca));
    signs of the indices are ignored, as in the reference (553)."""
    j = np.abs(uvidx4) - 1
    Vk = V[j]                                   # [4, Nca]
    amp = np.abs(Vk)
    model = np.log(amp[0]) + np.log(amp[1]) - np.log(amp[2]) - np.log(amp[3])
    r = ca - model
    chisq = np.sum(r ** 2 / var * fnorm)
    factor = -2.0 * r / var * fnorm
    for k, sg in ((0, 1.0), (1, 1.0), (2, -1.0), (3, -1.0)):
        f = sg * factor / amp[k] ** 2
        _scatter_add(Vresre, j[k], f * Vk[k].real)
        _scatter_add(Vresim, j[k], f * Vk[k].imag)
    return chisq


def chisq_cp(V, uvidx3, cp, var, fnorm, Vresre, Vresim):
    """fftlib.f90:576-655  (the 2(1-cos r) form, not the legacy wrapped r^2
    form of calc_chisq_cp_old)."""
    j = np.abs(uvidx3) - 1
    s = np.sign(uvidx3).astype(np.float64)
    Vk = V[j]                                   # [3, Ncp]
    ampsq = np.abs(Vk) ** 2
    model = np.sum(np.arctan2(Vk.imag, Vk.real) * s, axis=0)
    r = cp - model
    chisq = np.sum(2.0 * (1.0 - np.cos(r)) / var * fnorm)
    factor = -2.0 * np.sin(r) / var * fnorm
    for k in range(3):
        f = factor / ampsq[k] * s[k]
        _scatter_add(Vresre, j[k], -f * Vk[k].imag)
        _scatter_add(Vresim, j[k], f * Vk[k].real)
    return chisq


# ----------------------------------------------------------------------------
# Regularizers  (image.f90)
# ----------------------------------------------------------------------------
def init_l1reg(prior):
    """image.f90:207-223."""
    w = 1.0 / smabs(prior)
    return w / np.sum(smabs(prior) * w)


def init_tvreg(xidx, yidx, Nx, Ny, isflat, prior):
    """image.f90:409-446."""
    N = prior.size
    if isflat:
        return np.full(N, 1.0 / (ZEROEPS + np.sum(prior)) / N)
    I2d = image_1d_to_2d(prior, xidx, yidx, Nx, Ny)
    w = 1.0 / tv_e_map(I2d)[yidx - 1, xidx - 1]
    tv = np.sum(tv_e_map(I2d)[yidx - 1, xidx - 1] * w)
    return w / tv


def init_tsvreg(xidx, yidx, Nx, Ny, isflat, prior):
    """image.f90:620-655."""
    N = prior.size
    if isflat:
        return np.full(N, 1.0 / (ZEROEPS + np.sum(prior) ** 2 / N) / N)
    w = 1.0 / (prior ** 2 + ZEROEPS)
    I2d = image_1d_to_2d(prior, xidx, yidx, Nx, Ny)
    tsv = np.sum(tsv_e_map(I2d)[yidx - 1, xidx - 1] * w)
    return w / tsv


def kl_e(I, P):
    a = smabs(I)
    return a * np.log(a / P) + P / math.e


def kl_grade(I, P):
    a = smabs(I)
    return smabs_diff(I) * (np.log(a / P) + 1.0)


def init_klreg(kl_l_in, prior):
    """image.f90:278-291: weight = smabs(prior); lambda / kl(P, P)."""
    wgt = smabs(prior)
    return kl_l_in / np.sum(kl_e(wgt, wgt)), wgt


def gs_e(I, P):
    a = smabs(I)
    return a * (np.log(a / P) - 1.0) + P


def gs_grade(I, P):
    a = smabs(I)
    return smabs_diff(I) * np.log(a / P)


def init_gsreg(prior):
    """image.f90:352-364."""
    return smabs(prior)


def init_tfdreg(tfd_l_in, tgtfd, tgter):
    """image.f90:760-771."""
    return tfd_l_in / (tgtfd * tgter) ** 2


def image_1d_to_2d(I1d, xidx, yidx, Nx, Ny):
    """I1d_I2d_fwd, image.f90:1167-1186 (zero outside the window)."""
    I2d = np.zeros((Ny, Nx), dtype=np.float64)
    I2d[np.asarray(yidx) - 1, np.asarray(xidx) - 1] = I1d
    return I2d


def _shift(I2d, dy, dx):
    """out[j,i] = I2d[j+dy, i+dx], zero where out of range."""
    Ny, Nx = I2d.shape
    out = np.zeros_like(I2d)
    ys = slice(max(0, -dy), min(Ny, Ny - dy))
    xs = slice(max(0, -dx), min(Nx, Nx - dx))
    yd = slice(max(0, dy), min(Ny, Ny + dy))
    xd = slice(max(0, dx), min(Nx, Nx + dx))
    out[ys, xs] = I2d[yd, xd]
    return out


def _fwd_diffs(I2d):
    """dIx = I(i+1,j)-I(i,j) (0 at i=Nx), dIy = I(i,j+1)-I(i,j) (0 at j=Ny)."""
    Ny, Nx = I2d.shape
    dIx = np.zeros_like(I2d)
    dIy = np.zeros_like(I2d)
    dIx[:, :Nx - 1] = I2d[:, 1:] - I2d[:, :-1]
    dIy[:Ny - 1, :] = I2d[1:, :] - I2d[:-1, :]
    return dIx, dIy


def tv_e_map(I2d):
    """tv_e at every pixel (image.f90:477-508)."""
    dIx, dIy = _fwd_diffs(I2d)
    return np.sqrt(dIx * dIx + dIy * dIy)


def tsv_e_map(I2d):
    """tsv_e at every pixel (image.f90:678-709)."""
    dIx, dIy = _fwd_diffs(I2d)
    return dIx * dIx + dIy * dIy


def tv_grade_map(I2d):
    """tv_grade at every pixel, with the reference's reset quirk
    (image.f90:512-612): a neighbour term whose magnitude is <= zeroeps
    resets the accumulated gradient to 0."""
    Ny, Nx = I2d.shape
    dIx, dIy = _fwd_diffs(I2d)
    e = np.sqrt(dIx * dIx + dIy * dIy)
    g = np.where(e > ZEROEPS, -(dIx + dIy) / np.where(e > ZEROEPS, e, 1.0), 0.0)
    # term of pixel (i-1, j): valid where i > 1
    dIx2 = np.zeros_like(I2d)
    dIy2 = np.zeros_like(I2d)
    dIx2[:, 1:] = I2d[:, 1:] - I2d[:, :-1]
    dIy2[:Ny - 1, 1:] = I2d[1:, :-1] - I2d[:-1, :-1]
    e2 = np.sqrt(dIx2 * dIx2 + dIy2 * dIy2)
    valid2 = np.zeros_like(I2d, dtype=bool)
    valid2[:, 1:] = True
    g = np.where(valid2, np.where(e2 > ZEROEPS, g + dIx2 / np.where(e2 > ZEROEPS, e2, 1.0), 0.0), g)
    # term of pixel (i, j-1): valid where j > 1
    dIy3 = np.zeros_like(I2d)
    dIx3 = np.zeros_like(I2d)
    dIy3[1:, :] = I2d[1:, :] - I2d[:-1, :]
    dIx3[1:, :Nx - 1] = I2d[:-1, 1:] - I2d[:-1, :-1]
    e3 = np.sqrt(dIx3 * dIx3 + dIy3 * dIy3)
    valid3 = np.zeros_like(I2d, dtype=bool)
    valid3[1:, :] = True
    g = np.where(valid3, np.where(e3 > ZEROEPS, g + dIy3 / np.where(e3 > ZEROEPS, e3, 1.0), 0.0), g)
    return g


def tsv_grade_map(I2d):
    """tsv_grade at every pixel (image.f90:713-754)."""
    Ny, Nx = I2d.shape
    g = np.zeros_like(I2d)
    g[:, :Nx - 1] -= 2.0 * (I2d[:, 1:] - I2d[:, :-1])
    g[:Ny - 1, :] -= 2.0 * (I2d[1:, :] - I2d[:-1, :])
    g[:, 1:] += 2.0 * (I2d[:, 1:] - I2d[:, :-1])
    g[1:, :] += 2.0 * (I2d[1:, :] - I2d[:-1, :])
    return g


def calc_tfdreg(I1d, tgtfd):
    """image.f90:774-787: cost = (sum I - F)^2, gradient 2 (sum I - F)."""
    resid = np.sum(I1d) - tgtfd
    return resid ** 2, 2.0 * resid


def calc_cenreg(I1d, xidx, yidx, Nxref, Nyref, alpha):
    """image.f90:793-863."""
    dix = xidx - Nxref
    diy = yidx - Nyref
    if abs(alpha - 1.0) < ZEROEPS:
        Ip = smabs(I1d)
        gradIp = smabs_diff(I1d)
    else:
        Ip = smabs(I1d) ** alpha
        gradIp = smabs_diff(I1d) * alpha * smabs(I1d) ** (alpha - 1.0)
    sumIp = ZEROEPS + np.sum(Ip)
    sumx = np.sum(Ip * dix) / sumIp
    sumy = np.sum(Ip * diy) / sumIp
    cost = math.sqrt(sumx ** 2 + sumy ** 2 + ZEROEPS)
    grad = (sumx * (dix - sumx) + sumy * (diy - sumy)) / sumIp / cost * gradIp
    return cost, grad


def _sigma_prior(maj, mn, phi):
    """image.f90:1021-1031."""
    return np.array([mn * math.cos(phi) ** 2 + maj * math.sin(phi) ** 2,
                     mn * math.sin(phi) ** 2 + maj * math.cos(phi) ** 2,
                     (maj - mn) * math.sin(phi) * math.cos(phi)])


def calc_smreg(I1d, xidx, yidx, maj, mn, phi):
    """Second-moment regularizer: sm_e / sm_grade and helpers,
    image.f90:964-1138 (cost without lambda, per-pixel gradient)."""
    x = xidx.astype(np.float64)
    y = yidx.astype(np.float64)
    # totalflux / xy_cen use abs(), not smabs (image.f90:1108-1138)
    Itot = np.sum(np.abs(I1d))
    Isum = Itot
    xcen = np.sum(x * I1d) / Itot
    ycen = np.sum(y * I1d) / Itot
    Sg = np.array([np.sum(I1d * (x - xcen) ** 2 / Isum),
                   np.sum(I1d * (y - ycen) ** 2 / Isum),
                   np.sum(I1d * (x - xcen) * (y - ycen) / Isum)])
    mom = np.array([np.sum((x - xcen) * I1d), np.sum((y - ycen) * I1d)])
    Sp = _sigma_prior(maj, mn, phi)
    cost = (Sg[0] - Sp[0]) ** 2 + (Sg[1] - Sp[1]) ** 2 + 2.0 * (Sg[2] - Sp[2]) ** 2
    xg = (x - xcen) / Isum
    yg = (y - ycen) / Isum
    g1 = (((x - xcen) ** 2 - 2.0 * xg * mom[0]) - Sg[0]) / Isum
    g2 = (((y - ycen) ** 2 - 2.0 * yg * mom[1]) - Sg[1]) / Isum
    g3 = ((x - xcen) * (y - ycen) - xg * mom[1] - yg * mom[0] - Sg[2]) / Isum
    grad = 2.0 * (Sg[0] - Sp[0]) * g1 + 2.0 * (Sg[1] - Sp[1]) * g2 + 4.0 * (Sg[2] - Sp[2]) * g3
    return cost, grad


# ----------------------------------------------------------------------------
# Problem container and the full cost  (fftim2d.f90:387-607)
# ----------------------------------------------------------------------------
@dataclass
class Problem:
    """Arguments of calc_cost (fftim2d.f90:387-407) after the set-up of
    fftim2d.imaging (weights normalised, Vfcv phase-shifted)."""
    Nx: int
    Ny: int
    xidx: np.ndarray            # int, 1-based, [Npix]
    yidx: np.ndarray
    u: np.ndarray               # scaled u' [M]
    v: np.ndarray
    Nxref: float = 0.0
    Nyref: float = 0.0
    # regularizers: lambda <= 0 disables (image.f90:111-190)
    l1_l: float = -1.0
    l1_wgt: np.ndarray | None = None
    tv_l: float = -1.0
    tv_wgt: np.ndarray | None = None
    tsv_l: float = -1.0
    tsv_wgt: np.ndarray | None = None
    kl_l: float = -1.0
    kl_wgt: np.ndarray | None = None
    gs_l: float = -1.0
    gs_wgt: np.ndarray | None = None
    tfd_l: float = -1.0
    tfd_tgtfd: float = 1.0
    cen_l: float = -1.0
    cen_alpha: float = 1.0
    sm_l: float = -1.0
    sm_maj: float = 0.0
    sm_min: float = 0.0
    sm_phi: float = 0.0
    # data (already phase-shifted / weight-normalised)
    uvidxfcv: np.ndarray | None = None
    Vfcv: np.ndarray | None = None
    Varfcv: np.ndarray | None = None
    wfcv: float = 0.0
    uvidxamp: np.ndarray | None = None
    Vamp: np.ndarray | None = None
    Varamp: np.ndarray | None = None
    wamp: float = 0.0
    uvidxcp: np.ndarray | None = None     # [3, Ncp]
    CP: np.ndarray | None = None
    Varcp: np.ndarray | None = None
    wcp: float = 0.0
    uvidxca: np.ndarray | None = None     # [4, Nca]
    CA: np.ndarray | None = None
    Varca: np.ndarray | None = None
    wca: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def npix(self):
        return int(np.asarray(self.xidx).size)

    @property
    def isfcv(self):
        return self.uvidxfcv is not None and len(self.uvidxfcv) > 0

    @property
    def isamp(self):
        return self.uvidxamp is not None and len(self.uvidxamp) > 0

    @property
    def iscp(self):
        return self.uvidxcp is not None and self.uvidxcp.shape[-1] > 0

    @property
    def isca(self):
        return self.uvidxca is not None and self.uvidxca.shape[-1] > 0


TERM_NAMES = ("chisq", "chisq_fcv", "chisq_amp", "chisq_cp", "chisq_ca",
              "reg", "l1", "tv", "tsv", "kl", "gs", "tfd", "cen", "sm", "cost")


def calc_chisq(x, p: Problem, fwd=None, adj=None):
    """fftlib.f90:250-382.  Returns (chisq, gradchisq[Npix], per-term dict,
    V, Vresre, Vresim).  fwd/adj default to the exact direct DFT."""
    fwd = fwd or dft_fwd
    I2d = image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny)
    V = fwd(I2d, p.u, p.v)
    chisq, t, Vresre, Vresim = _chisq_blocks(V, p)
    g2d = nufft_adj_resid(Vresre, Vresim, p.u, p.v, p.Nx, p.Ny, adj)
    grad = g2d[np.asarray(p.yidx) - 1, np.asarray(p.xidx) - 1]
    return chisq, grad, t, V, Vresre, Vresim


def _chisq_blocks(V, p: Problem):
    """The four data blocks of calc_chisq on model visibilities V
    (fftlib.f90:338-367; the same sequence in fftlib3d.f90:96-121)."""
    M = V.size
    Vresre = np.zeros(M)
    Vresim = np.zeros(M)
    chisq = 0.0
    t = dict(chisq_fcv=0.0, chisq_amp=0.0, chisq_cp=0.0, chisq_ca=0.0)
    if p.isfcv:
        c = chisq_fcv(V, p.uvidxfcv, p.Vfcv, p.Varfcv, p.wfcv, Vresre, Vresim)
        chisq += c
        t["chisq_fcv"] = c / p.wfcv / (2 * len(p.uvidxfcv))
    if p.isamp:
        c = chisq_amp(V, p.uvidxamp, p.Vamp, p.Varamp, p.wamp, Vresre, Vresim)
        chisq += c
        t["chisq_amp"] = c / p.wamp / len(p.uvidxamp)
    if p.isca:
        c = chisq_ca(V, p.uvidxca, p.CA, p.Varca, p.wca, Vresre, Vresim)
        chisq += c
        t["chisq_ca"] = c / p.wca / p.uvidxca.shape[1]
    if p.iscp:
        c = chisq_cp(V, p.uvidxcp, p.CP, p.Varcp, p.wcp, Vresre, Vresim)
        chisq += c
        t["chisq_cp"] = c / p.wcp / p.uvidxcp.shape[1]
    t["chisq"] = chisq
    return chisq, t, Vresre, Vresim


def calc_cost_reg(x, p: Problem):
    """image.f90:12-200.  Returns (reg, gradreg[Npix], per-term dict)."""
    N = x.size
    xidx = np.asarray(p.xidx)
    yidx = np.asarray(p.yidx)
    g = np.zeros(N)
    t = dict(l1=0.0, tv=0.0, tsv=0.0, kl=0.0, gs=0.0, tfd=0.0, cen=0.0, sm=0.0)
    if p.tfd_l > 0:
        c, gg = calc_tfdreg(x, p.tfd_tgtfd)
        t["tfd"] = p.tfd_l * c
        g += p.tfd_l * gg
    if p.cen_l > 0:
        c, gg = calc_cenreg(x, xidx, yidx, p.Nxref, p.Nyref, p.cen_alpha)
        t["cen"] = p.cen_l * c
        g += p.cen_l * gg
    if p.sm_l > 0:
        c, gg = calc_smreg(x, xidx, yidx, p.sm_maj, p.sm_min, p.sm_phi)
        t["sm"] = p.sm_l * c
        g += p.sm_l * gg
    if p.l1_l > 0:
        t["l1"] = np.sum(p.l1_l * p.l1_wgt * smabs(x))
        g += p.l1_l * p.l1_wgt * smabs_diff(x)
    if p.kl_l > 0:
        t["kl"] = np.sum(p.kl_l * kl_e(x, p.kl_wgt))
        g += p.kl_l * kl_grade(x, p.kl_wgt)
    if p.gs_l > 0:
        t["gs"] = np.sum(p.gs_l * gs_e(x, p.gs_wgt))
        g += p.gs_l * gs_grade(x, p.gs_wgt)
    if p.tv_l > 0 or p.tsv_l > 0:
        I2d = image_1d_to_2d(x, xidx, yidx, p.Nx, p.Ny)
        sel = (yidx - 1, xidx - 1)
        if p.tv_l > 0:
            t["tv"] = np.sum(p.tv_l * p.tv_wgt * tv_e_map(I2d)[sel])
            g += p.tv_l * p.tv_wgt * tv_grade_map(I2d)[sel]
        if p.tsv_l > 0:
            t["tsv"] = np.sum(p.tsv_l * p.tsv_wgt * tsv_e_map(I2d)[sel])
            g += p.tsv_l * p.tsv_wgt * tsv_grade_map(I2d)[sel]
    reg = sum(t.values())
    return reg, g, t


def calc_cost(x, p: Problem, fwd=None, adj=None):
    """fftim2d.f90:387-607: cost = chisq + reg, grad = gradchisq + gradreg.
    Returns (cost, grad, terms) with terms keyed by TERM_NAMES."""
    x = np.asarray(x, dtype=np.float64)
    chisq, gchi, tc, *_ = calc_chisq(x, p, fwd, adj)
    reg, greg, tr = calc_cost_reg(x, p)
    terms = dict(tc)
    terms.update(tr)
    terms["reg"] = reg
    terms["cost"] = chisq + reg
    return chisq + reg, gchi + greg, terms


def terms_vector(terms):
    return np.array([terms[k] for k in TERM_NAMES], dtype=np.float64)


# ----------------------------------------------------------------------------
# dynamic imaging: fftim3d.calc_cost (fftim3d.f90:472-640)
# ----------------------------------------------------------------------------
TERM_NAMES_MOVIE = TERM_NAMES + ("rt", "ri", "lc")


@dataclass
class Movie:
    """Arguments of fftim3d.calc_cost (fftim3d.f90:472-498): ``frame`` holds
    the geometry, the static regularizers and the data over the
    concatenation of the frames' uv lists (frame z owns Nuvs[z] consecutive
    points, tools.py get_uvlist_loop 255-336).  Image vector: Nz * Npix,
    frame-major.  lc_l None = light-curve mode off (lbfgs3d.py:528-530);
    lc_tgtfd None = ones."""
    frame: Problem
    Nuvs: np.ndarray
    rt_l: float = -1.0
    rt_wgt: np.ndarray | None = None
    ri_l: float = -1.0
    ri_wgt: np.ndarray | None = None
    lc_l: np.ndarray | None = None
    lc_tgtfd: np.ndarray | None = None
    lc_nidx: int = -1
    extra: dict = field(default_factory=dict)

    @property
    def Nz(self):
        return int(np.asarray(self.Nuvs).size)

    @property
    def lc(self):
        return np.ones(self.Nz) if self.lc_tgtfd is None else np.asarray(self.lc_tgtfd, np.float64)

    @property
    def lc_mode(self):
        return self.lc_l is not None and float(np.sum(self.lc_l)) > 0


def calc_chisq3d(x, mv: Movie, fwd=None, adj=None):
    """fftlib3d.f90:14-201: per-frame type-2 transforms into one
    visibility vector, the shared chi-square blocks, per-frame adjoints of
    each frame's residual slice.  Frames without uv points are skipped
    (fftlib3d.f90:77, 127).  Returns (chisq, grad[Nz*Npix], terms, V)."""
    p = mv.frame
    fwd = fwd or dft_fwd
    Nz, npix = mv.Nz, p.npix
    nuvs = np.asarray(mv.Nuvs, np.int64)
    start = np.concatenate([[0], np.cumsum(nuvs)])
    V = np.zeros(p.u.size, np.complex128)
    for z in range(Nz):
        if nuvs[z] == 0:
            continue
        s, e = start[z], start[z + 1]
        I2d = image_1d_to_2d(x[z * npix:(z + 1) * npix], p.xidx, p.yidx, p.Nx, p.Ny)
        V[s:e] = fwd(I2d, p.u[s:e], p.v[s:e])
    chisq, t, Vresre, Vresim = _chisq_blocks(V, p)
    grad = np.zeros(Nz * npix)
    sel = (np.asarray(p.yidx) - 1, np.asarray(p.xidx) - 1)
    for z in range(Nz):
        if nuvs[z] == 0:
            continue
        s, e = start[z], start[z + 1]
        g2d = nufft_adj_resid(Vresre[s:e], Vresim[s:e], p.u[s:e], p.v[s:e], p.Nx, p.Ny, adj)
        grad[z * npix:(z + 1) * npix] = g2d[sel]
    return chisq, grad, t, V


def frame_problem(mv: Movie, z):
    """The 2-D regularizer arguments frame z sees in calc_cost_reg3d
    (image3d.f90:184-236): centroid off; in light-curve mode the frame's
    lc_l / lc_tgtfd replace tfd_l / tfd_tgtfd and, for lc_nidx 1 or 3, the
    weights are rescaled by tfd_tgtfd / lc_tgtfd (squared for TSV)."""
    p = mv.frame
    rep = dict(cen_l=-1.0, cen_alpha=1.0)
    if mv.lc_mode:
        lc = mv.lc
        rep.update(tfd_l=float(mv.lc_l[z]), tfd_tgtfd=float(lc[z]))
        if mv.lc_nidx in (1, 3):
            s = p.tfd_tgtfd / lc[z]
            for name in ("l1", "tv", "kl", "gs"):
                w = getattr(p, name + "_wgt")
                if w is not None:
                    rep[name + "_wgt"] = np.asarray(w) * p.tfd_tgtfd / lc[z]
            if p.tsv_wgt is not None:
                rep["tsv_wgt"] = np.asarray(p.tsv_wgt) * s ** 2
    return dataclasses.replace(p, **rep)


def calc_cost_dynamical(x, mv: Movie):
    """image3d.f90:280-392 (Rt and Ri; Rs, Rf and the centroid term stay 0
    there).  Returns (rt_cost, ri_cost, grad[Nz*Npix])."""
    Nz, npix = mv.Nz, mv.frame.npix
    lc = mv.lc
    Iin = np.asarray(x, np.float64).reshape(Nz, npix)
    Iavg = np.zeros(npix)
    for z in range(Nz):
        Iavg = Iavg + (Iin[z] / lc[z]) / Nz
    rt_cost = 0.0
    ri_cost = 0.0
    g = np.zeros((Nz, npix))
    for z in range(Nz):
        II = Iin[z] / lc[z]
        Iu = Iin[z + 1] / lc[z + 1] if z < Nz - 1 else np.zeros(npix)
        Il = Iin[z - 1] / lc[z - 1] if z > 0 else np.zeros(npix)
        if mv.rt_l > 0 and z < Nz - 1:
            rt_cost += np.sum(mv.rt_l * mv.rt_wgt * (Iu - II) ** 2)
            # rt_grade (image3d.f90:409-421); its iz == Nz branch is never
            # reached because of the z < Nz - 1 guard above
            gr = 2.0 * (II - Iu) if z == 0 else 2.0 * (2.0 * II - Il - Iu)
            g[z] += mv.rt_l * mv.rt_wgt * gr / lc[z]
        if mv.ri_l > 0:
            ri_cost += np.sum(mv.ri_l * mv.ri_wgt * (II - Iavg) ** 2)
            g[z] += mv.ri_l * mv.ri_wgt * (2.0 * (II - Iavg)) / lc[z]
    return rt_cost, ri_cost, g.reshape(-1)


def calc_cost_reg3d(x, mv: Movie):
    """image3d.f90:15-274.  Returns (reg, gradreg[Nz*Npix], terms)."""
    Nz, npix = mv.Nz, mv.frame.npix
    x = np.asarray(x, np.float64)
    g = np.zeros(Nz * npix)
    keys = ("l1", "tv", "tsv", "kl", "gs", "tfd", "cen", "sm")
    t = {k: 0.0 for k in keys}
    t.update(rt=0.0, ri=0.0, lc=0.0)
    reg = 0.0
    for z in range(Nz):
        pz = frame_problem(mv, z)
        rz, gz, tz = calc_cost_reg(x[z * npix:(z + 1) * npix], pz)
        for k in keys:
            if k == "tfd" and mv.lc_mode:
                t["lc"] += tz[k] / Nz
            else:
                t[k] += tz[k] / Nz
        reg += rz / Nz
        g[z * npix:(z + 1) * npix] = gz / Nz
    if mv.frame.cen_l > 0 or mv.rt_l > 0 or mv.ri_l > 0:
        rt, ri, gd = calc_cost_dynamical(x, mv)
        t["rt"], t["ri"] = rt, ri
        reg += rt + ri
        g = g + gd
    return reg, g, t


def calc_cost3d(x, mv: Movie, fwd=None, adj=None):
    """fftim3d.f90:472-640: cost = chisq3d + reg3d.  Returns (cost, grad,
    terms) with terms keyed by TERM_NAMES_MOVIE."""
    x = np.asarray(x, np.float64)
    chisq, gchi, tc, _ = calc_chisq3d(x, mv, fwd, adj)
    reg, greg, tr = calc_cost_reg3d(x, mv)
    terms = dict(tc)
    terms.update(tr)
    terms["reg"] = reg
    terms["cost"] = chisq + reg
    return chisq + reg, gchi + greg, terms


def terms_vector_movie(terms):
    return np.array([terms[k] for k in TERM_NAMES_MOVIE], dtype=np.float64)


# ----------------------------------------------------------------------------
# uv list and problem set-up (tools.py:12-132, fftim2d.f90:178-241)
# ----------------------------------------------------------------------------
def get_uvlist(fcv_uv=None, amp_uv=None, cp_uv=None, ca_uv=None):
    """tools.py:12-132 on plain arrays.

    fcv_uv / amp_uv: (u, v) arrays; cp_uv: ((u12,v12),(u23,v23),(u31,v31));
    ca_uv: ((u1,v1),...,(u4,v4)).  Returns (u, v, uvidxfcv, uvidxamp,
    uvidxcp[3,Ncp], uvidxca[4,Nca]) with signed 1-based indices.
    """
    stacks, signs = [], []

    def add(u, v):
        u = np.array(u, dtype=np.float64)
        v = np.array(v, dtype=np.float64)
        sg = np.ones(u.size, dtype=np.int64)
        neg = v < 0
        u[neg] *= -1
        v[neg] *= -1
        sg[neg] = -1
        stacks.append(u + 1j * v)
        signs.append(sg)

    nf = na = nc = nq = 0
    if fcv_uv is not None:
        add(*fcv_uv)
        nf = len(fcv_uv[0])
    if amp_uv is not None:
        add(*amp_uv)
        na = len(amp_uv[0])
    if cp_uv is not None:
        for uv in cp_uv:
            add(*uv)
        nc = len(cp_uv[0][0])
    if ca_uv is not None:
        for uv in ca_uv:
            add(*uv)
        nq = len(ca_uv[0][0])
    uvstack = np.concatenate(stacks)
    sgstack = np.concatenate(signs)
    uniq, inv = np.unique(uvstack, return_inverse=True)
    uvidx = (inv.reshape(-1) + 1) * sgstack
    o = 0
    fcv = uvidx[o:o + nf] if nf else np.zeros(0, np.int64)
    o += nf
    amp = uvidx[o:o + na] if na else np.zeros(0, np.int64)
    o += na
    cp = uvidx[o:o + 3 * nc].reshape([nc, 3], order="F").T if nc else np.zeros((3, 0), np.int64)
    o += 3 * nc
    ca = uvidx[o:o + 4 * nq].reshape([nq, 4], order="F").T if nq else np.zeros((4, 0), np.int64)
    return (uniq.real.copy(), uniq.imag.copy(), fcv.astype(np.int32), amp.astype(np.int32),
            cp.astype(np.int32), ca.astype(np.int32))


def normalise_weights(nfcv, wfcv, namp, wamp, ncp, wcp, nca, wca):
    """fftim2d.f90:178-213 -> (wfcv_n, wamp_n, wcp_n, wca_n)."""
    wtotal = 2 * nfcv * wfcv + namp * wamp + ncp * wcp + nca * wca
    return (wfcv / wtotal if nfcv else 0.0, wamp / wtotal if namp else 0.0,
            wcp / wtotal if ncp else 0.0, wca / wtotal if nca else 0.0)


def phashift_r2c(Vfcv, uvidxfcv, u, v, Nxref, Nyref, Nx, Ny):
    """fftim2d.f90:224-241 with phashift_r2c (fftlib.f90:223-244): the signed
    uv of each datum, dix = Nxref - Nx/2 - 1."""
    s = np.sign(uvidxfcv)
    j = np.abs(uvidxfcv) - 1
    ut = u[j] * s
    vt = v[j] * s
    dix = Nxref - Nx / 2.0 - 1.0
    diy = Nyref - Ny / 2.0 - 1.0
    return Vfcv * np.exp(1j * (ut * dix + vt * diy))
