"""ctypes binding of oracle/_ref/libsmili_ref.so: the reference's own
regularizer module (src/fortlib/image.f90 with src/fortlib/param.f90),
compiled from the reference sources by `make -C oracle ref` -- test
infrastructure only, used to pin the numpy oracle's regularizers and
initialisers.  `load()` returns None where the library was not built (the
reference tree is not everywhere: the tests that need it skip there)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libsmili_ref.so")
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int32)
_lib = None


def load():
    global _lib
    if _lib is None and os.path.exists(LIB):
        lib = C.CDLL(LIB)
        d, i = C.c_double, C.c_int32
        lib.ref_calc_cost_reg.argtypes = [i, i, i, _pd, _pi, _pi, d, d, d, _pd, d, _pd, d, _pd, d, _pd, d, _pd,
                                          d, d, d, d, d, d, d, d, _pd, _pd, _pd]
        lib.ref_init_l1reg.argtypes = [i, _pd, _pd]
        lib.ref_init_tvreg.argtypes = [i, i, i, _pi, _pi, i, _pd, _pd]
        lib.ref_init_tsvreg.argtypes = [i, i, i, _pi, _pi, i, _pd, _pd]
        lib.ref_init_klreg.argtypes = [i, d, _pd, _pd, _pd]
        lib.ref_init_gsreg.argtypes = [i, _pd, _pd]
        lib.ref_init_tfdreg.argtypes = [d, d, d, _pd]
        lib.ref_calc_cenreg.argtypes = [i, _pd, _pi, _pi, d, d, d, _pd, _pd]
        for f in ("ref_calc_cost_reg", "ref_init_l1reg", "ref_init_tvreg", "ref_init_tsvreg", "ref_init_klreg",
                  "ref_init_gsreg", "ref_init_tfdreg", "ref_calc_cenreg"):
            getattr(lib, f).restype = None
        _lib = lib
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_pd)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_pi)


REG_NAMES = ("l1", "tv", "tsv", "kl", "gs", "tfd", "cen", "sm")


def calc_cost_reg(x, xidx, yidx, Nx, Ny, Nxref=0.0, Nyref=0.0, **reg):
    """image.calc_cost_reg: (cost, grad, {term: cost}, (out_maj, out_min,
    out_phi)).  reg holds <term>_l and <term>_wgt for l1 .. gs and the tfd /
    cen / sm parameters, with the reference's defaults (lambda -1 = off)."""
    lib = load()
    n = x.size
    ones = np.ones(n)
    keep = []

    def dp(a):
        arr, ptr = _d(a)
        keep.append(arr)
        return ptr

    x_, xp = _d(x)
    xi, xip = _i(xidx)
    yi, yip = _i(yidx)
    costs = np.zeros(11)
    cost = np.zeros(1)
    grad = np.zeros(n)
    args = [n, Nx, Ny, xp, xip, yip, Nxref, Nyref]
    for r in ("l1", "tv", "tsv", "kl", "gs"):
        w = reg.get(r + "_wgt")
        args += [reg.get(r + "_l", -1.0), dp(ones if w is None else w)]
    args += [reg.get("tfd_l", -1.0), reg.get("tfd_tgtfd", 1.0), reg.get("cen_l", -1.0),
             reg.get("cen_alpha", 1.0), reg.get("sm_l", -1.0), reg.get("sm_maj", 1.0),
             reg.get("sm_min", 1.0), reg.get("sm_phi", 0.0)]
    args += [costs.ctypes.data_as(_pd), cost.ctypes.data_as(_pd), grad.ctypes.data_as(_pd)]
    lib.ref_calc_cost_reg(*args)
    return float(cost[0]), grad, dict(zip(REG_NAMES, costs[:8].tolist())), tuple(costs[8:].tolist())


def init_l1reg(prior):
    p, pp = _d(prior)
    w = np.zeros(p.size)
    load().ref_init_l1reg(p.size, pp, w.ctypes.data_as(_pd))
    return w


def init_tvreg(xidx, yidx, Nx, Ny, isflat, prior):
    p, pp = _d(prior)
    xi, xip = _i(xidx)
    yi, yip = _i(yidx)
    w = np.zeros(p.size)
    load().ref_init_tvreg(p.size, Nx, Ny, xip, yip, int(bool(isflat)), pp, w.ctypes.data_as(_pd))
    return w


def init_tsvreg(xidx, yidx, Nx, Ny, isflat, prior):
    p, pp = _d(prior)
    xi, xip = _i(xidx)
    yi, yip = _i(yidx)
    w = np.zeros(p.size)
    load().ref_init_tsvreg(p.size, Nx, Ny, xip, yip, int(bool(isflat)), pp, w.ctypes.data_as(_pd))
    return w


def init_klreg(l_in, prior):
    p, pp = _d(prior)
    w = np.zeros(p.size)
    lo = np.zeros(1)
    load().ref_init_klreg(p.size, l_in, pp, lo.ctypes.data_as(_pd), w.ctypes.data_as(_pd))
    return float(lo[0]), w


def init_gsreg(prior):
    p, pp = _d(prior)
    w = np.zeros(p.size)
    load().ref_init_gsreg(p.size, pp, w.ctypes.data_as(_pd))
    return w


def init_tfdreg(l_in, tgtfd, tgter):
    lo = np.zeros(1)
    load().ref_init_tfdreg(l_in, tgtfd, tgter, lo.ctypes.data_as(_pd))
    return float(lo[0])


def calc_cenreg(x, xidx, yidx, Nxref, Nyref, alpha):
    x_, xp = _d(x)
    xi, xip = _i(xidx)
    yi, yip = _i(yidx)
    c = np.zeros(1)
    g = np.zeros(x_.size)
    load().ref_calc_cenreg(x_.size, xp, xip, yip, Nxref, Nyref, alpha, c.ctypes.data_as(_pd),
                           g.ctypes.data_as(_pd))
    return float(c[0]), g
