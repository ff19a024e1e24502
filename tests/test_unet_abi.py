"""C ABI of the warm-start network that needs no GPU: libsmili_unet.so
loads, exports every entry point include/smili_unet.h declares, agrees with
the oracle on the parameter count, and rejects invalid arguments before any
HIP call."""
import ctypes as C
import os
import re

import pytest

from oracle import unet_oracle as U
from smili_amd.warmstart import unet

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "smili_unet.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smili_unet_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = unet.load()
    names = header_functions()
    assert len(names) == 14
    for n in names:
        assert hasattr(lib, n), n


def test_parameter_count_matches_oracle():
    assert unet.nparams() == U.nparams() == 1946593


@pytest.mark.parametrize("b,h,w", [(0, 128, 128), (1, 32, 128), (1, 128, 64), (1, 100, 128), (5000, 64, 128)])
def test_invalid_sizes_rejected(b, h, w):
    lib = unet.load()
    out = C.c_void_p()
    assert lib.smili_unet_create(b, h, w, C.byref(out)) == -1
    assert not out.value
    assert lib.smili_unet_last_error()


def test_null_arguments():
    lib = unet.load()
    assert lib.smili_unet_create(1, 64, 128, None) == -1
    assert lib.smili_unet_set_params(None, None) == -1
    assert lib.smili_unet_forward(None, None, 1, None) == -1
    assert lib.smili_unet_loss_grad(None, None, None, 1, None, None) == -1
    assert lib.smili_unet_train(None, None, 1, None, None, None) == -1
    assert lib.smili_unet_flops(None, None) == -1
    lib.smili_unet_destroy(None)


def test_wrapper_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(unet, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(unet, "_lib", None)
    with pytest.raises(unet.UNetError):
        unet.nparams()


def test_packed_weights_round_trip():
    """The packed weights form keeps the first convolution, biases and head
    in float32 and rounds the other 3x3 weights to bfloat16 exactly as the
    kernels do."""
    import numpy as np
    rng = np.random.default_rng(0)
    p = U.init_params(rng)
    p[U.nparams() - 40:] = rng.standard_normal(40)
    q = unet.unpack(**unet.pack(p))
    assert q.dtype == np.float32 and q.shape == p.shape
    ref = p.copy()
    o = 0
    for i, (_, ci, co) in enumerate(U.layer_specs()):
        nw = co * 9 * ci
        if i > 0:
            ref[o:o + nw] = U.bf16(p[o:o + nw])
        o += nw + co
    assert np.array_equal(q, ref)
    assert [(n, ci, co) for n, ci, co in unet.LAYERS] == U.layer_specs()

