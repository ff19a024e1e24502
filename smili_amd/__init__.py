"""MI355X-native (gfx950) RML imaging cost of SMILI.

The hot path -- objective + gradient of the regularized maximum-likelihood
imaging cost (src/fortlib/fftim2d.f90:387-607 in the reference) -- runs as
hand-written HIP kernels in ``libsmili_hip.so`` behind a C ABI
(``include/smili_hip.h``); Python reaches it through ctypes.
"""
from ._lib import NTERMS, TERM_NAMES, SmiliError, load  # noqa: F401

__all__ = ["NTERMS", "TERM_NAMES", "SmiliError", "load"]
