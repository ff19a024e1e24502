"""ctypes wrapper of oracle/libsmili_cpu.so, the C + OpenMP FP64 restatement
of the cost (test infrastructure / CPU baseline only)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsmili_cpu.so")


class CpuCost:
    def __init__(self, **kw):
        from smili_amd import _lib
        from smili_amd.cost import build_desc
        if not os.path.exists(LIB):
            raise OSError(LIB + " not built (make -C oracle)")
        self.lib = C.CDLL(LIB)
        self.lib.smili_cpu_create.argtypes = [C.POINTER(_lib.SmiliDesc), C.POINTER(C.c_void_p)]
        self.lib.smili_cpu_cost_grad.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
        self.lib.smili_cpu_destroy.argtypes = [C.c_void_p]
        self.lib.smili_cpu_threads.argtypes = [C.c_void_p]
        desc, keep = build_desc(**{k: v for k, v in kw.items() if k != "extra"})
        self.npix = int(desc.npix)
        h = C.c_void_p()
        if self.lib.smili_cpu_create(C.byref(desc), C.byref(h)) != 0:
            raise RuntimeError("smili_cpu_create failed")
        self.h = h
        del keep

    @property
    def threads(self):
        return self.lib.smili_cpu_threads(self.h)

    def cost_grad(self, x):
        x = np.ascontiguousarray(x, np.float64)
        g = np.empty(self.npix)
        t = np.empty(15)
        c = C.c_double()
        dp = C.POINTER(C.c_double)
        self.lib.smili_cpu_cost_grad(self.h, x.ctypes.data_as(dp), C.byref(c), g.ctypes.data_as(dp),
                                     t.ctypes.data_as(dp))
        return c.value, g, t

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.smili_cpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
