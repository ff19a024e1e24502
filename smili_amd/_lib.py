"""ctypes binding of libsmili_hip.so (C ABI declared in include/smili_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``python -m smili_amd.build``).  There is no fallback: if the library cannot
be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmili_hip.so")

NTERMS = 15
TERM_NAMES = ("chisq", "chisq_fcv", "chisq_amp", "chisq_cp", "chisq_ca",
              "reg", "l1", "tv", "tsv", "kl", "gs", "tfd", "cen", "sm", "cost")
NTERMS_MOVIE = 18
TERM_NAMES_MOVIE = TERM_NAMES + ("rt", "ri", "lc")

SMILI_OK = 0
SMILI_ENONFINITE = -4

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int32)


class SmiliReg(C.Structure):
    _fields_ = [("l1_l", C.c_double), ("l1_wgt", _pd),
                ("tv_l", C.c_double), ("tv_wgt", _pd),
                ("tsv_l", C.c_double), ("tsv_wgt", _pd),
                ("kl_l", C.c_double), ("kl_wgt", _pd),
                ("gs_l", C.c_double), ("gs_wgt", _pd),
                ("tfd_l", C.c_double), ("tfd_tgtfd", C.c_double),
                ("cen_l", C.c_double), ("cen_alpha", C.c_double),
                ("sm_l", C.c_double), ("sm_maj", C.c_double),
                ("sm_min", C.c_double), ("sm_phi", C.c_double)]


class SmiliData(C.Structure):
    _fields_ = [("nfcv", C.c_int32), ("uvidxfcv", _pi),
                ("vfcv_re", _pd), ("vfcv_im", _pd), ("varfcv", _pd), ("wfcv", C.c_double),
                ("namp", C.c_int32), ("uvidxamp", _pi),
                ("vamp", _pd), ("varamp", _pd), ("wamp", C.c_double),
                ("ncp", C.c_int32), ("uvidxcp", _pi),
                ("cp", _pd), ("varcp", _pd), ("wcp", C.c_double),
                ("nca", C.c_int32), ("uvidxca", _pi),
                ("ca", _pd), ("varca", _pd), ("wca", C.c_double)]


class SmiliDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32),
                ("nxref", C.c_double), ("nyref", C.c_double),
                ("npix", C.c_int32), ("xidx", _pi), ("yidx", _pi),
                ("nuv", C.c_int32), ("u", _pd), ("v", _pd),
                ("reg", SmiliReg), ("data", SmiliData)]


class SmiliMovieDesc(C.Structure):
    _fields_ = [("frame", SmiliDesc), ("nz", C.c_int32), ("nuvs", _pi),
                ("rt_l", C.c_double), ("rt_wgt", _pd),
                ("ri_l", C.c_double), ("ri_wgt", _pd),
                ("lc_l", _pd), ("lc_tgtfd", _pd), ("lc_nidx", C.c_int32)]


class SmiliTune(C.Structure):
    _fields_ = [("bin", C.c_int32), ("sub_points", C.c_int32),
                ("col_group", C.c_int32), ("chunk", C.c_int32),
                ("use_graph", C.c_int32), ("precision", C.c_int32), ("branches", C.c_int32),
                ("fft_radix", C.c_int32), ("chisq_threads", C.c_int32), ("transform", C.c_int32),
                ("deterministic", C.c_int32), ("dft_shapes", C.c_int32)]


# every symbol include/smili_hip.h declares, with (restype, argtypes)
SIGNATURES = {
    "smili_abi_version": (C.c_int, []),
    "smili_last_error": (C.c_char_p, []),
    "smili_problem_create": (C.c_int, [C.POINTER(SmiliDesc), C.POINTER(SmiliTune), C.POINTER(C.c_void_p)]),
    "smili_problem_destroy": (None, [C.c_void_p]),
    "smili_cost_grad": (C.c_int, [C.c_void_p, _pd, _pd, _pd, _pd]),
    "smili_cost_grad_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "smili_cost_grad_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "smili_synchronize": (C.c_int, [C.c_void_p]),
    "smili_problem_buffers": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p)]),
    "smili_eval_async": (C.c_int, [C.c_void_p]),
    "smili_problem_host_buffers": (C.c_int, [C.c_void_p, C.POINTER(_pd), C.POINTER(_pd)]),
    "smili_problem_stream": (C.c_void_p, [C.c_void_p]),
    "smili_model_vis": (C.c_int, [C.c_void_p, _pd, _pd, _pd]),
    "smili_adjoint_resid": (C.c_int, [C.c_void_p, _pd, _pd, _pd]),
    "smili_problem_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "smili_problem_unit_info": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]),
    "smili_problem_profile": (C.c_int, [C.c_void_p, C.c_int32, _pd]),
    "smili_batch_create": (C.c_int, [C.POINTER(SmiliDesc), C.c_int32, C.POINTER(SmiliTune),
                                     C.POINTER(C.c_void_p)]),
    "smili_batch_cost_grad": (C.c_int, [C.c_void_p, _pd, _pd, _pd, _pd]),
    "smili_movie_create": (C.c_int, [C.POINTER(SmiliMovieDesc), C.POINTER(SmiliTune), C.POINTER(C.c_void_p)]),
    "smili_problem_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int32)]),
    "smili_transform_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _pd, _pd, C.POINTER(SmiliTune),
                                         C.POINTER(C.c_void_p)]),
}

STAGE_NAMES = ("scatter", "t2_rows", "t2_cols", "interp", "moments", "moments_reduce",
               "chisq", "spread", "t1_cols", "t1_rows", "reg", "finalize", "frame_mean",
               "dft2_y", "dft2_x", "dft1_x", "dft1_y", "gather")
KERNEL_OF_STAGE = dict(zip(STAGE_NAMES, ("k_scatter", "k_t2_rows", "k_t2_cols", "k_interp", "k_moments",
                                         "k_moments_reduce", "k_chisq", "k_spread", "k_t1_cols",
                                         "k_t1_rows", "k_reg", "k_finalize", "k_frame_mean",
                                         "k_dft", "k_dft", "k_dft", "k_dft", "k_spread_gather")))

_lib = None


class SmiliError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"smili_hip error {code}: {msg}")
        self.code = code


def load(path: str | None = None):
    """Load libsmili_hip.so (raises OSError if it is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.smili_abi_version() != 3:
        raise OSError("libsmili_hip ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def check(rc, allow_nonfinite=False):
    if rc == SMILI_OK or (allow_nonfinite and rc == SMILI_ENONFINITE):
        return rc
    msg = load().smili_last_error()
    raise SmiliError(rc, msg.decode() if msg else "")


def dptr(a):
    return a.ctypes.data_as(_pd) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_pi) if a is not None else None
