"""Host wrappers of the imaging costs on the GPU.

``CostProblem`` takes the argument list of the reference's
``fftim2d.calc_cost`` (src/fortlib/fftim2d.f90:387-407) -- image geometry,
scaled uv list, regularizer lambdas and weights, and the four data blocks
with signed 1-based uv indices -- uploads it once, and then evaluates
``cost, grad, terms = problem.cost_grad(x)`` on the GPU through the C ABI of
libsmili_hip.so.  Keyword names follow calc_cost's argument names.

``MovieProblem`` does the same for the dynamic-imaging cost
``fftim3d.calc_cost`` (src/fortlib/fftim3d.f90:472-640) and
``BatchProblem`` for many independent 2-D problems evaluated together.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import NTERMS, NTERMS_MOVIE, TERM_NAMES, TERM_NAMES_MOVIE, check, dptr, iptr


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def build_desc(Nx, Ny, xidx, yidx, u, v, Nxref=0.0, Nyref=0.0,
               l1_l=-1.0, l1_wgt=None, tv_l=-1.0, tv_wgt=None,
               tsv_l=-1.0, tsv_wgt=None, kl_l=-1.0, kl_wgt=None,
               gs_l=-1.0, gs_wgt=None, tfd_l=-1.0, tfd_tgtfd=1.0,
               cen_l=-1.0, cen_alpha=1.0, sm_l=-1.0, sm_maj=0.0, sm_min=0.0, sm_phi=0.0,
               uvidxfcv=None, Vfcv=None, Varfcv=None, wfcv=0.0,
               uvidxamp=None, Vamp=None, Varamp=None, wamp=0.0,
               uvidxcp=None, CP=None, Varcp=None, wcp=0.0,
               uvidxca=None, CA=None, Varca=None, wca=0.0, extra=None):
    """smili_desc for calc_cost's argument list; returns (desc, keep) where
    `keep` holds the arrays the descriptor points to (keep it alive until
    the descriptor has been consumed)."""
    keep = []

    def k(a):
        keep.append(a)
        return a

    xidx = k(_i32(xidx))
    yidx = k(_i32(yidx))
    u = k(_f64(u))
    v = k(_f64(v))
    npix = int(xidx.size)
    reg = _lib.SmiliReg()
    for name, lam, w in (("l1", l1_l, l1_wgt), ("tv", tv_l, tv_wgt), ("tsv", tsv_l, tsv_wgt),
                         ("kl", kl_l, kl_wgt), ("gs", gs_l, gs_wgt)):
        setattr(reg, name + "_l", float(lam))
        if lam > 0:
            arr = k(_f64(np.broadcast_to(np.asarray(w, np.float64), (npix,))))
            setattr(reg, name + "_wgt", dptr(arr))
    reg.tfd_l, reg.tfd_tgtfd = float(tfd_l), float(tfd_tgtfd)
    reg.cen_l, reg.cen_alpha = float(cen_l), float(cen_alpha)
    reg.sm_l, reg.sm_maj, reg.sm_min, reg.sm_phi = map(float, (sm_l, sm_maj, sm_min, sm_phi))
    data = _lib.SmiliData()
    if uvidxfcv is not None and len(uvidxfcv) > 0:
        Vfcv = np.asarray(Vfcv, np.complex128)
        data.nfcv = len(uvidxfcv)
        data.uvidxfcv = iptr(k(_i32(uvidxfcv)))
        data.vfcv_re = dptr(k(_f64(Vfcv.real)))
        data.vfcv_im = dptr(k(_f64(Vfcv.imag)))
        data.varfcv = dptr(k(_f64(Varfcv)))
        data.wfcv = float(wfcv)
    if uvidxamp is not None and len(uvidxamp) > 0:
        data.namp = len(uvidxamp)
        data.uvidxamp = iptr(k(_i32(uvidxamp)))
        data.vamp = dptr(k(_f64(Vamp)))
        data.varamp = dptr(k(_f64(Varamp)))
        data.wamp = float(wamp)
    if uvidxcp is not None and np.asarray(uvidxcp).shape[-1] > 0:
        idx = np.asarray(uvidxcp).reshape(3, -1)
        data.ncp = idx.shape[1]
        data.uvidxcp = iptr(k(_i32(idx.T)))          # 3 contiguous per closure
        data.cp = dptr(k(_f64(CP)))
        data.varcp = dptr(k(_f64(Varcp)))
        data.wcp = float(wcp)
    if uvidxca is not None and np.asarray(uvidxca).shape[-1] > 0:
        idx = np.asarray(uvidxca).reshape(4, -1)
        data.nca = idx.shape[1]
        data.uvidxca = iptr(k(_i32(idx.T)))          # 4 contiguous per closure
        data.ca = dptr(k(_f64(CA)))
        data.varca = dptr(k(_f64(Varca)))
        data.wca = float(wca)
    desc = _lib.SmiliDesc()
    desc.nx, desc.ny = int(Nx), int(Ny)
    desc.nxref, desc.nyref = float(Nxref), float(Nyref)
    desc.npix = npix
    desc.xidx, desc.yidx = iptr(xidx), iptr(yidx)
    desc.nuv = int(u.size)
    desc.u, desc.v = dptr(u), dptr(v)
    desc.reg = reg
    desc.data = data
    return desc, keep


def _tune(tune):
    if not tune:
        return None
    t = _lib.SmiliTune()
    for k, val in tune.items():
        setattr(t, k, int(val))
    return t


class _Handle:
    """Common part of the handle wrappers: lifetime, device buffers, the
    transforms and profiling."""

    _term_names = TERM_NAMES

    def _adopt(self, lib, h):
        self._lib = lib
        self._h = h
        nprob, nxt, nterms = C.c_int32(), C.c_int64(), C.c_int32()
        check(lib.smili_problem_sizes(h, C.byref(nprob), C.byref(nxt), C.byref(nterms)))
        self.nprob, self.nx_total, self.nterms = int(nprob.value), int(nxt.value), int(nterms.value)

    # ------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._hostbufs = None
            self._lib.smili_problem_destroy(self._h)
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

    # ------------------------------------------------------------------
    def cost_grad_device(self, x_ptr, grad_ptr, terms_ptr=None):
        """Asynchronous evaluation on device pointers (ints)."""
        check(self._lib.smili_cost_grad_device(self._h, C.c_void_p(x_ptr), C.c_void_p(grad_ptr),
                                                C.c_void_p(terms_ptr) if terms_ptr else None))

    def cost_grad_device_noalloc(self, x_ptr, grad_ptr, terms_ptr):
        """smili_cost_grad_device_noalloc: the evaluation with its
        intermediates in the library's static device scratch (no
        allocation); device pointers, asynchronous on the handle's stream."""
        check(self._lib.smili_cost_grad_device_noalloc(self._h, C.c_void_p(x_ptr), C.c_void_p(grad_ptr),
                                                       C.c_void_p(terms_ptr)))

    def buffers(self):
        """Device addresses (ints) of the handle's own x, grad and terms."""
        x, g, t = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(self._lib.smili_problem_buffers(self._h, C.byref(x), C.byref(g), C.byref(t)))
        return x.value, g.value, t.value

    def host_buffers(self):
        """The handle's page-locked host x and grad (numpy views, valid
        until close()): evaluating with x in the first and grad into the
        second (``cost_grad(x_view, out=grad_view)``) skips the host copies
        of the host-pointer path."""
        if getattr(self, "_hostbufs", None) is None:
            x, g = _lib._pd(), _lib._pd()
            check(self._lib.smili_problem_host_buffers(self._h, C.byref(x), C.byref(g)))
            n = self.nx_total
            self._hostbufs = (np.ctypeslib.as_array(x, shape=(n,)), np.ctypeslib.as_array(g, shape=(n,)))
        return self._hostbufs

    def eval_async(self):
        """One evaluation on the handle's own buffers (no copies, async)."""
        check(self._lib.smili_eval_async(self._h))

    def synchronize(self):
        check(self._lib.smili_synchronize(self._h))

    @property
    def stream(self):
        return self._lib.smili_problem_stream(self._h)

    def model_vis(self, x):
        """Complex model visibilities at the uv points (type-2 transforms),
        all frames / problems in order."""
        x = _f64(x)
        if x.size != self.nx_total:
            raise ValueError(f"x has {x.size} entries, expected {self.nx_total}")
        vre = np.empty(self.nuv)
        vim = np.empty(self.nuv)
        check(self._lib.smili_model_vis(self._h, dptr(x), dptr(vre), dptr(vim)))
        return vre + 1j * vim

    def adjoint_resid(self, vres_re, vres_im):
        """Re(A1 (vres_re - i vres_im)) at the window pixels of every frame /
        problem."""
        out = np.empty(self.nx_total)
        check(self._lib.smili_adjoint_resid(self._h, dptr(_f64(vres_re)), dptr(_f64(vres_im)), dptr(out)))
        return out

    def profile(self, reps=10, back_to_back=False):
        """Average device ms per kernel stage (eager launches with hipEvents);
        back_to_back: one pass with each stage launched `reps` times between
        its events (per-launch time with the event overhead amortised; the
        handle's grad and terms are stale afterwards)."""
        ms = np.zeros(len(_lib.STAGE_NAMES))
        check(self._lib.smili_problem_profile(self._h, -int(reps) if back_to_back else int(reps), dptr(ms)))
        return dict(zip(_lib.STAGE_NAMES, ms.tolist()))

    def unit_info(self, unit=0):
        """Transform and grid of one unit (problem of a batch / movie frame):
        dict(transform=0 none | 1 fft | 2 dft, gw, gh, m, nx, ny, nf1, nf2)."""
        a = (C.c_int64 * 8)()
        check(self._lib.smili_problem_unit_info(self._h, int(unit), a))
        return dict(zip(("transform", "gw", "gh", "m", "nx", "ny", "nf1", "nf2"), list(a)))

    def info(self):
        a = (C.c_int64 * 8)()
        check(self._lib.smili_problem_info(self._h, a))
        keys = ("nf1", "nf2", "nuv", "subproblems", "bin", "chunks", "launches", "graph")
        return dict(zip(keys, list(a)))

    def _cost_grad(self, x, allow_nonfinite, out=None):
        x = _f64(x)
        if x.size != self.nx_total:
            raise ValueError(f"x has {x.size} entries, expected {self.nx_total}")
        if out is not None and (out.dtype != np.float64 or out.size != self.nx_total or
                                not out.flags.c_contiguous):
            raise ValueError("out must be a contiguous float64 array of the image-vector length")
        grad = np.empty(self.nx_total) if out is None else out
        terms = np.empty(self.nterms)
        cost = C.c_double()
        check(self._lib.smili_cost_grad(self._h, dptr(x), C.byref(cost), dptr(grad), dptr(terms)),
              allow_nonfinite)
        return cost.value, grad, dict(zip(self._term_names, terms.tolist()))


class CostProblem(_Handle):
    """One RML imaging objective resident on the current GPU.

    Parameters follow calc_cost (fftim2d.f90:387-407): ``xidx, yidx`` 1-based
    pixel coordinates of the npix image vector; ``u, v`` scaled unique uv
    points; ``uvidxcp`` of shape (3, Ncp) and ``uvidxca`` of shape (4, Nca)
    as the Fortran arrays uvidxcp(3,Ncp), uvidxca(4,Nca); Vfcv complex and
    already phase-shifted; w* normalised weights.  A regularizer is on iff
    its lambda > 0; a data block is on iff it is given and non-empty.
    ``tune`` sets smili_tune fields (bin, sub_points, col_group, chunk,
    use_graph, precision, branches, fft_radix, chisq_threads).
    """

    def __init__(self, tune=None, **kw):
        lib = _lib.load()
        desc, keep = build_desc(**kw)
        self.Nx, self.Ny = int(desc.nx), int(desc.ny)
        self.npix = int(desc.npix)
        self.nuv = int(desc.nuv)
        t = _tune(tune)
        h = C.c_void_p()
        check(lib.smili_problem_create(C.byref(desc), C.byref(t) if t is not None else None, C.byref(h)))
        del keep                  # the library copied everything it needs
        self._adopt(lib, h)

    def cost_grad(self, x, allow_nonfinite=False, out=None):
        """(cost, grad[npix], terms dict) for image vector x; ``out``: a
        float64 array of npix entries to write the gradient into (e.g. the
        handle's page-locked grad, host_buffers())."""
        return self._cost_grad(x, allow_nonfinite, out)


class MovieProblem(_Handle):
    """The dynamic-imaging objective of fftim3d.calc_cost
    (src/fortlib/fftim3d.f90:472-640) resident on the current GPU.

    ``Nz`` frames share the geometry and the static regularizers of the 2-D
    keywords (as CostProblem); ``u, v`` is the concatenation of the frames'
    uv lists with ``Nuvs[z]`` points for frame z, and the data's uv indices
    refer to that concatenation (tools.py get_uvlist_loop, 255-336).
    Temporal terms: ``rt_l, rt_wgt`` (Rt), ``ri_l, ri_wgt`` (Ri); light curve
    ``lc_l, lc_tgtfd, lc_nidx`` as calc_cost_reg3d (image3d.f90:15-274).
    The image vector is frame-major (Nz * npix).
    """

    _term_names = TERM_NAMES_MOVIE

    def __init__(self, Nz, Nuvs, rt_l=-1.0, rt_wgt=None, ri_l=-1.0, ri_wgt=None,
                 lc_l=None, lc_tgtfd=None, lc_nidx=-1, tune=None, **kw):
        lib = _lib.load()
        desc, keep = build_desc(**kw)
        self.Nx, self.Ny = int(desc.nx), int(desc.ny)
        self.npix = int(desc.npix)
        self.nuv = int(desc.nuv)
        self.Nz = int(Nz)
        md = _lib.SmiliMovieDesc()
        md.frame = desc
        md.nz = self.Nz
        nuvs = _i32(Nuvs)
        if nuvs.size != self.Nz:
            raise ValueError("Nuvs must have Nz entries")
        keep.append(nuvs)
        md.nuvs = iptr(nuvs)
        md.rt_l, md.ri_l = float(rt_l), float(ri_l)
        for name, lam, w in (("rt_wgt", rt_l, rt_wgt), ("ri_wgt", ri_l, ri_wgt)):
            if lam > 0:
                arr = _f64(np.broadcast_to(np.asarray(w, np.float64), (self.npix,)))
                keep.append(arr)
                setattr(md, name, dptr(arr))
        for name, arr in (("lc_l", lc_l), ("lc_tgtfd", lc_tgtfd)):
            if arr is not None:
                arr = _f64(np.broadcast_to(np.asarray(arr, np.float64), (self.Nz,)))
                keep.append(arr)
                setattr(md, name, dptr(arr))
        md.lc_nidx = int(lc_nidx)
        t = _tune(tune)
        h = C.c_void_p()
        check(lib.smili_movie_create(C.byref(md), C.byref(t) if t is not None else None, C.byref(h)))
        del keep
        self._adopt(lib, h)

    def cost_grad(self, x, allow_nonfinite=False, out=None):
        """(cost, grad[Nz * npix], terms dict incl. rt, ri, lc); ``out`` as
        for CostProblem.cost_grad."""
        return self._cost_grad(x, allow_nonfinite, out)


class BatchProblem(_Handle):
    """Many independent 2-D objectives evaluated together (one launch per
    kernel stage for all of them).  ``problems`` is a list of keyword dicts
    as CostProblem takes them."""

    def __init__(self, problems, tune=None):
        lib = _lib.load()
        descs = (_lib.SmiliDesc * len(problems))()
        keep = []
        self.npixs, self.nuvs = [], []
        for i, kw in enumerate(problems):
            d, k = build_desc(**kw)
            keep.append(k)
            descs[i] = d
            self.npixs.append(int(d.npix))
            self.nuvs.append(int(d.nuv))
        self.nuv = int(sum(self.nuvs))
        self.offsets = np.concatenate([[0], np.cumsum(self.npixs)]).astype(np.int64)
        t = _tune(tune)
        h = C.c_void_p()
        check(lib.smili_batch_create(descs, len(problems), C.byref(t) if t is not None else None, C.byref(h)))
        del keep
        self._adopt(lib, h)

    def split(self, vec):
        """Per-problem views of a concatenated image vector."""
        return [vec[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def cost_grad(self, x, allow_nonfinite=False, out=None):
        """(costs[nprob], grad (concatenated), terms[nprob, 15]) for the
        concatenated image vectors x."""
        x = _f64(x)
        if x.size != self.nx_total:
            raise ValueError(f"x has {x.size} entries, expected {self.nx_total}")
        grad = np.empty(self.nx_total)
        terms = np.empty((self.nprob, NTERMS))
        costs = np.empty(self.nprob)
        check(self._lib.smili_batch_cost_grad(self._h, dptr(x), dptr(costs), dptr(grad), dptr(terms)),
              allow_nonfinite)
        return costs, grad, terms
