"""GPU counterparts of the reference's transform and model routines, with
f2py's calling conventions (``fortlib.fftlib.*`` / ``fortlib.fftlib3d.*``):

* the NUFFT wrappers of src/fortlib/fftlib.f90:70-198 (FINUFFT type 2 with
  +i and type 1 with +i, src/wrapfinufft/wrapfinufft.f90:94-290) -- image
  arrays are indexed ``I2d[ix, iy]`` (Fortran ``I2d(Nx,Ny)``, as the callers
  pass ``np.asfortranarray(image.T)``, smili/uvdata/uvfits/uvfits.py:2612-2616),
  1-D images are x fastest;
* ``phashift_c2r`` / ``phashift_r2c`` (fftlib.f90:200-244);
* ``model_fcv / model_amp / model_ca / model_cp`` (fftlib.f90:744-1254) and
  their movie versions (src/fortlib/fftlib3d.f90:207-940): model and
  residual per datum, the un-normalised chi-square and its gradient.

Transforms, chi-square and gradient run on the GPU through the C ABI
(a transform handle, or a cost handle holding only the one data block with
weight 1, which makes calc_cost's chi-square and gradient exactly the
model_* ones); the per-datum model / residual values are elementwise
functions of the GPU model visibilities.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .cost import CostProblem, MovieProblem, _Handle, _f64, _tune
from ._lib import check, dptr


class Nufft(_Handle):
    """Transform handle on an Nx x Ny image and nuv scaled uv points."""

    def __init__(self, u, v, Nx, Ny, tune=None):
        lib = _lib.load()
        u = _f64(u)
        v = _f64(v)
        if u.shape != v.shape or u.ndim != 1:
            raise ValueError("u and v must be 1-D of equal length")
        self.Nx, self.Ny, self.nuv = int(Nx), int(Ny), int(u.size)
        t = _tune(tune)
        h = C.c_void_p()
        check(lib.smili_transform_create(self.Nx, self.Ny, self.nuv, dptr(u), dptr(v),
                                         C.byref(t) if t is not None else None, C.byref(h)))
        self._adopt(lib, h)

    def fwd(self, I2d):
        """finufft2d2 with +i of a real image I2d[ix, iy]."""
        x = np.ascontiguousarray(np.asarray(I2d, np.float64).T).reshape(-1)   # x fastest
        return self.model_vis(x)

    def adj_re(self, a, b):
        """Real part of A1 (a + i b), x fastest."""
        return self.adjoint_resid(a, -np.asarray(b, np.float64))

    def adj_im(self, a, b):
        """Imaginary part of A1 (a + i b), x fastest."""
        return self.adjoint_resid(b, a)


def _xy(flat, Nx, Ny):
    return np.asarray(flat).reshape(Ny, Nx).T          # -> [ix, iy]


def nufft_fwd(u, v, I2d):
    """fftlib.f90:70-84 -> Vcmp."""
    I2d = np.asarray(I2d, np.float64)
    with Nufft(u, v, I2d.shape[0], I2d.shape[1]) as n:
        return n.fwd(I2d)


def nufft_fwd_real(u, v, I2d):
    """fftlib.f90:86-100 -> (Vreal, Vimag)."""
    V = nufft_fwd(u, v, I2d)
    return V.real.copy(), V.imag.copy()


def nufft_adj(u, v, Vcmp, Nx, Ny):
    """fftlib.f90:102-116: complex I2d[ix, iy] = A1 Vcmp."""
    Vcmp = np.asarray(Vcmp, np.complex128)
    with Nufft(u, v, Nx, Ny) as n:
        re = n.adj_re(Vcmp.real, Vcmp.imag)
        im = n.adj_im(Vcmp.real, Vcmp.imag)
    return _xy(re, Nx, Ny) + 1j * _xy(im, Nx, Ny)


def nufft_adj_real1d(u, v, Vreal, Vimag, Nx, Ny):
    """fftlib.f90:118-132: Re(A1 (Vreal + i Vimag)) as an Nx*Ny vector (x
    fastest)."""
    with Nufft(u, v, Nx, Ny) as n:
        return n.adj_re(Vreal, Vimag)


def nufft_adj_real(u, v, Vreal, Vimag, Nx, Ny):
    """fftlib.f90:134-150 -> (Ireal, Iimag), Nx*Ny vectors."""
    with Nufft(u, v, Nx, Ny) as n:
        return n.adj_re(Vreal, Vimag), n.adj_im(Vreal, Vimag)


def nufft_adj_resid(u, v, Vre, Vim, Nx, Ny):
    """fftlib.f90:152-174: Re(A1 Vre) + Im(A1 Vim) for complex Vre, Vim."""
    Vre = np.asarray(Vre, np.complex128)
    Vim = np.asarray(Vim, np.complex128)
    with Nufft(u, v, Nx, Ny) as n:
        out = n.adj_re(Vre.real, Vre.imag) + n.adj_im(Vim.real, Vim.imag)
    return _xy(out, Nx, Ny)


def nufft_adjrea(u, v, Vcmp, Nx, Ny):
    """fftlib.f90:177-198: Re(A1 Vcmp) as I2d[ix, iy]."""
    Vcmp = np.asarray(Vcmp, np.complex128)
    with Nufft(u, v, Nx, Ny) as n:
        return _xy(n.adj_re(Vcmp.real, Vcmp.imag), Nx, Ny)


def phashift_c2r(u, v, Nxref, Nyref, Nx, Ny, Vcmp):
    """fftlib.f90:200-221: image-centre phase -> reference-pixel phase."""
    dix = Nx / 2.0 + 1 - Nxref
    diy = Ny / 2.0 + 1 - Nyref
    return np.asarray(Vcmp) * np.exp(1j * (np.asarray(u) * dix + np.asarray(v) * diy))


def phashift_r2c(u, v, Nxref, Nyref, Nx, Ny, Vcmp):
    """fftlib.f90:223-244: reference-pixel phase -> image-centre phase."""
    dix = Nxref - Nx / 2.0 - 1
    diy = Nyref - Ny / 2.0 - 1
    return np.asarray(Vcmp) * np.exp(1j * (np.asarray(u) * dix + np.asarray(v) * diy))


# ----------------------------------------------------------------------------
# model_* (fftlib.f90:744-1254, fftlib3d.f90:207-940)
# ----------------------------------------------------------------------------
def _signed_uv(u, v, idx):
    j = np.abs(idx) - 1
    s = np.sign(idx)
    return np.asarray(u)[j] * s, np.asarray(v)[j] * s


def _pick(V, idx):
    j = np.abs(idx) - 1
    return np.where(idx > 0, V[j], np.conj(V[j]))


def _run(geom, movie, block, tune):
    """chi-square (weight 1), gradient and model visibilities of one data
    block on the GPU."""
    kw = _strip(geom)
    kw.update(block)
    if movie is None:
        h = CostProblem(tune=tune, **kw)
    else:
        h = MovieProblem(tune=tune, Nz=movie[0], Nuvs=movie[1], **kw)
    with h:
        x = geom["_x"]
        cost, grad, terms = h.cost_grad(x)
        V = h.model_vis(x)
    return terms["chisq"], grad, V


def _geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v):
    return dict(Nx=int(Nx), Ny=int(Ny), xidx=np.asarray(xidx, np.int32), yidx=np.asarray(yidx, np.int32),
                Nxref=float(Nxref), Nyref=float(Nyref), u=np.asarray(u, np.float64),
                v=np.asarray(v, np.float64), _x=np.asarray(Iin, np.float64))


def _strip(g):
    return {k: v for k, v in g.items() if k != "_x"}


def _model_fcv(g, movie, uvidxfcv, Vfcvr, Vfcvi, Varfcv, tune):
    idx = np.asarray(uvidxfcv, np.int32).reshape(-1)
    us, vs = _signed_uv(g["u"], g["v"], idx)
    Vobs = phashift_r2c(us, vs, g["Nxref"], g["Nyref"], g["Nx"], g["Ny"],
                        np.asarray(Vfcvr, np.float64) + 1j * np.asarray(Vfcvi, np.float64))
    chisq, grad, V = _run(g, movie,
                          dict(uvidxfcv=idx, Vfcv=Vobs, Varfcv=Varfcv, wfcv=1.0), tune)
    model = _pick(V, idx)
    resid = Vobs - model
    resid = phashift_c2r(us, vs, g["Nxref"], g["Nyref"], g["Nx"], g["Ny"], resid)
    model = phashift_c2r(us, vs, g["Nxref"], g["Nyref"], g["Nx"], g["Ny"], model)
    return chisq, grad, model.real.copy(), model.imag.copy(), resid.real.copy(), resid.imag.copy()


def _model_amp(g, movie, uvidxamp, Vamp, Varamp, tune):
    idx = np.asarray(uvidxamp, np.int32).reshape(-1)
    chisq, grad, V = _run(g, movie, dict(uvidxamp=idx, Vamp=Vamp, Varamp=Varamp, wamp=1.0), tune)
    model = np.abs(V[np.abs(idx) - 1])
    return chisq, grad, model, np.asarray(Vamp, np.float64) - model


def _model_ca(g, movie, uvidxca, CA, Varca, tune):
    idx = np.asarray(uvidxca, np.int32).reshape(4, -1)
    chisq, grad, V = _run(g, movie, dict(uvidxca=idx, CA=CA, Varca=Varca, wca=1.0), tune)
    a = [np.log(np.abs(V[np.abs(idx[k]) - 1])) for k in range(4)]
    model = a[0] + a[1] - a[2] - a[3]
    return chisq, grad, model, np.asarray(CA, np.float64) - model


def _model_cp(g, movie, uvidxcp, CP, Varcp, tune):
    idx = np.asarray(uvidxcp, np.int32).reshape(3, -1)
    chisq, grad, V = _run(g, movie, dict(uvidxcp=idx, CP=CP, Varcp=Varcp, wcp=1.0), tune)
    model = sum(np.arctan2(V[np.abs(idx[k]) - 1].imag, V[np.abs(idx[k]) - 1].real) * np.sign(idx[k])
                for k in range(3))
    resid = np.asarray(CP, np.float64) - model
    resid = np.arctan2(np.sin(resid), np.cos(resid))
    model = np.arctan2(np.sin(model), np.cos(model))
    return chisq, grad, model, resid


def model_fcv(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v, uvidxfcv, Vfcvr, Vfcvi, Varfcv, tune=None):
    """fftlib.f90:744-896 -> (chisq, gradchisq, modelr, modeli, residr,
    residi); data in the reference-pixel phase, outputs likewise."""
    return _model_fcv(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), None,
                      uvidxfcv, Vfcvr, Vfcvi, Varfcv, tune)


def model_amp(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v, uvidxamp, Vamp, Varamp, tune=None):
    """fftlib.f90:898-997 -> (chisq, gradchisq, model, resid)."""
    return _model_amp(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), None, uvidxamp, Vamp, Varamp, tune)


def model_ca(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v, uvidxca, CA, Varca, tune=None):
    """fftlib.f90:999-1124 -> (chisq, gradchisq, model, resid)."""
    return _model_ca(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), None, uvidxca, CA, Varca, tune)


def model_cp(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v, uvidxcp, CP, Varcp, tune=None):
    """fftlib.f90:1126-1254 (closure-phase residuals wrapped to (-pi, pi])
    -> (chisq, gradchisq, model, resid)."""
    return _model_cp(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), None, uvidxcp, CP, Varcp, tune)


class fftlib3d:
    """Movie versions (src/fortlib/fftlib3d.f90:207-940): Iin is Nz frames
    (frame-major), u, v the concatenated per-frame uv lists with Nuvs[z]
    points for frame z."""

    @staticmethod
    def model_fcv(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, Nz, u, v, Nuvs, uvidxfcv, Vfcvr, Vfcvi, Varfcv,
                  tune=None):
        return _model_fcv(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), (Nz, Nuvs),
                          uvidxfcv, Vfcvr, Vfcvi, Varfcv, tune)

    @staticmethod
    def model_amp(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, Nz, u, v, Nuvs, uvidxamp, Vamp, Varamp, tune=None):
        return _model_amp(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), (Nz, Nuvs),
                          uvidxamp, Vamp, Varamp, tune)

    @staticmethod
    def model_ca(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, Nz, u, v, Nuvs, uvidxca, CA, Varca, tune=None):
        return _model_ca(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), (Nz, Nuvs), uvidxca, CA, Varca,
                         tune)

    @staticmethod
    def model_cp(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, Nz, u, v, Nuvs, uvidxcp, CP, Varcp, tune=None):
        return _model_cp(_geom(Iin, xidx, yidx, Nxref, Nyref, Nx, Ny, u, v), (Nz, Nuvs), uvidxcp, CP, Varcp,
                         tune)
