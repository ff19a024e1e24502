"""ctypes binding of libsmili_lbfgsb.so (include/smili_lbfgsb.h): the native
FP64 L-BFGS-B that stands in for the reference's src/lbfgsb/lbfgsb.f.

``Lbfgsb`` is one problem's optimizer driven by reverse communication, as
the reference's loops drive setulb (fftim2d.f90:286-351):

    opt = Lbfgsb(n, m=5, lower=0.0)        # nonnegative, as imaging(nonneg=True)
    task = opt.step(x)                      # START: x projected, task FG
    while task in (FG, NEW_X):
        if task == FG:
            f, g = fg(x)
        task = opt.step(x, f, g)

``minimize(fg, x0, ...)`` runs that loop; ``BatchLbfgsb`` steps many
independent problems at once over host threads.  The library is built
in-tree by ``smili_amd.build``; there is no fallback when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmili_lbfgsb.so")

FG, NEW_X, CONVERGENCE, ABNORMAL, ERROR = 1, 2, 3, 4, 5
STOPPED_BY_CALLBACK = "STOP: CALLBACK"
ITERATION_LIMIT = "STOP: TOTAL NO. of ITERATIONS REACHED LIMIT"

_lib = None
_pd = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -m smili_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.smili_lbfgsb_create.restype = C.c_void_p
        L.smili_lbfgsb_create.argtypes = [C.c_int64, C.c_int32, _pd, _pd, C.POINTER(C.c_int32),
                                          C.c_double, C.c_double, C.c_int32, C.c_int32]
        L.smili_lbfgsb_destroy.argtypes = [C.c_void_p]
        L.smili_lbfgsb_step.restype = C.c_int32
        L.smili_lbfgsb_step.argtypes = [C.c_void_p, _pd, C.c_double, _pd]
        L.smili_lbfgsb_step_batch.restype = C.c_int32
        L.smili_lbfgsb_step_batch.argtypes = [C.c_int32, C.POINTER(C.c_void_p), C.POINTER(_pd), _pd,
                                              C.POINTER(_pd), C.POINTER(C.c_int32), C.c_int32]
        L.smili_lbfgsb_info.restype = C.c_int32
        L.smili_lbfgsb_info.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_char_p,
                                        C.c_int32]
        _lib = L
    return _lib


def _bounds(n, lower, upper):
    """(l, u, nbd) arrays from scalar or per-variable bounds (None: none)."""
    lo = np.zeros(n) if lower is None else np.broadcast_to(np.asarray(lower, np.float64), (n,)).copy()
    up = np.zeros(n) if upper is None else np.broadcast_to(np.asarray(upper, np.float64), (n,)).copy()
    has_l = np.zeros(n, bool) if lower is None else np.isfinite(lo)
    has_u = np.zeros(n, bool) if upper is None else np.isfinite(up)
    nbd = np.where(has_l & has_u, 2, np.where(has_l, 1, np.where(has_u, 3, 0))).astype(np.int32)
    lo[~has_l] = 0.0
    up[~has_u] = 0.0
    return lo, up, nbd


class Lbfgsb:
    """One problem: n variables, m correction pairs, bounds lower/upper
    (scalars or arrays; None or +-inf: none), factr, pgtol and maxls as
    setulb's.  nthreads: host threads for the O(n m) work (0: all)."""

    def __init__(self, n, m=5, lower=None, upper=None, factr=1e7, pgtol=1e-5, maxls=20, nthreads=0):
        self.n = int(n)
        self.lo, self.up, self.nbd = _bounds(self.n, lower, upper)
        h = lib().smili_lbfgsb_create(self.n, int(m), self.lo.ctypes.data_as(_pd), self.up.ctypes.data_as(_pd),
                                      self.nbd.ctypes.data_as(C.POINTER(C.c_int32)), float(factr), float(pgtol),
                                      int(maxls), int(nthreads))
        if not h:
            raise ValueError("invalid L-BFGS-B set-up (n, m, maxls > 0; nbd in 0..3; l <= u)")
        self._h = C.c_void_p(h)

    def step(self, x, f=0.0, g=None):
        """One call of the reverse-communication protocol; x (float64,
        contiguous, n) is updated in place.  Returns the task."""
        if x.dtype != np.float64 or not x.flags.c_contiguous or x.size != self.n:
            raise ValueError("x must be a contiguous float64 array of n values")
        gp = None
        if g is not None:
            g = np.ascontiguousarray(g, dtype=np.float64)
            if g.size != self.n:
                raise ValueError("g must have n values")
            gp = g.ctypes.data_as(_pd)
        return int(lib().smili_lbfgsb_step(self._h, x.ctypes.data_as(_pd), float(f), gp))

    def info(self):
        """(iterations, evaluations, message)."""
        it, nf = C.c_int64(), C.c_int64()
        buf = C.create_string_buffer(128)
        lib().smili_lbfgsb_info(self._h, C.byref(it), C.byref(nf), buf, 128)
        return int(it.value), int(nf.value), buf.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            lib().smili_lbfgsb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def minimize(fg, x0, m=5, lower=None, upper=None, factr=1e7, pgtol=1e-5, maxiter=15000, maxls=20,
             callback=None, nthreads=0, xbuf=None):
    """L-BFGS-B from x0 with fg(x) -> (f, g).  callback(x) after every
    iteration (raise StopIteration to stop).  Stops after maxiter
    iterations.  xbuf: an array of n float64 to keep the iterate in.
    Returns (x, iterations, message, converged)."""
    x0 = np.asarray(x0, np.float64).ravel()
    x = xbuf if xbuf is not None else np.empty_like(x0)
    x[:] = x0
    with Lbfgsb(x0.size, m, lower, upper, factr, pgtol, maxls, nthreads) as opt:
        task = opt.step(x)
        nit = 0
        msg = None
        while task in (FG, NEW_X):
            if task == FG:
                f, g = fg(x)
                task = opt.step(x, f, g)
                continue
            nit += 1
            if callback is not None:
                xv = x.view()
                xv.flags.writeable = False
                try:
                    callback(xv)
                except StopIteration:
                    msg = STOPPED_BY_CALLBACK
                    break
            if nit >= maxiter:
                msg = ITERATION_LIMIT
                break
            task = opt.step(x)
        if msg is None:
            msg = opt.info()[2]
        return np.array(x, copy=True), nit, msg, task == CONVERGENCE


class BatchLbfgsb:
    """Independent problems stepped together: ``step(xs, fs, gs, active)``
    makes one reverse-communication call of every active problem (all by
    default; problem k reads fs[k], gs[k] when its last task was FG),
    spread over nthreads host threads.  xs: list of float64 arrays (one per
    problem, updated in place).  Returns the tasks (inactive problems keep
    their last one)."""

    def __init__(self, sizes, m=5, lower=None, upper=None, factr=1e7, pgtol=1e-5, maxls=20, nthreads=0):
        self.opts = [Lbfgsb(n, m, lower, upper, factr, pgtol, maxls, 1) for n in sizes]
        self.nthreads = int(nthreads)
        self.tasks = np.zeros(len(self.opts), np.int32)

    def step(self, xs, fs=None, gs=None, active=None):
        idx = [k for k in range(len(self.opts)) if active is None or active[k]]
        k = len(idx)
        if k == 0:
            return self.tasks.copy()
        hs = (C.c_void_p * k)(*[self.opts[i]._h.value for i in idx])
        for i in idx:
            if xs[i].dtype != np.float64 or not xs[i].flags.c_contiguous or xs[i].size != self.opts[i].n:
                raise ValueError("each x must be a contiguous float64 array of its problem's n values")
        xp = (_pd * k)(*[xs[i].ctypes.data_as(_pd) for i in idx])
        fa = np.array([0.0 if fs is None else float(fs[i]) for i in idx])
        gk = [np.zeros(self.opts[i].n) if gs is None else np.ascontiguousarray(gs[i], np.float64) for i in idx]
        gp = (_pd * k)(*[g.ctypes.data_as(_pd) for g in gk])         # gk keeps the arrays alive
        out = np.zeros(k, np.int32)
        lib().smili_lbfgsb_step_batch(k, hs, xp, fa.ctypes.data_as(_pd), gp,
                                      out.ctypes.data_as(C.POINTER(C.c_int32)), self.nthreads)
        self.tasks[idx] = out
        return self.tasks.copy()

    def close(self):
        for o in self.opts:
            o.close()
