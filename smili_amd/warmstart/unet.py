"""ctypes binding of libsmili_unet.so (C ABI in include/smili_unet.h).

The network, its parameter layout and its rounding points are defined in
include/smili_unet.h.  There is no fallback: without the library every
call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_HERE, "libsmili_unet.so")
_lib = None

_pf = C.POINTER(C.c_float)
_pd = C.POINTER(C.c_double)
_pl = C.POINTER(C.c_int64)


class UNetError(RuntimeError):
    pass


# (name, input channels, output channels) of the 3x3 convolutions in
# parameter order (include/smili_unet.h); then the head w[32], b[1]
LAYERS = (("conv_in", 2, 32), ("enc0b", 32, 32), ("enc1a", 32, 64), ("enc1b", 64, 64),
          ("enc2a", 64, 128), ("enc2b", 128, 128), ("enc3a", 128, 256), ("enc3b", 256, 256),
          ("dec2a", 384, 128), ("dec2b", 128, 128), ("dec1a", 192, 64), ("dec1b", 64, 64),
          ("dec0a", 96, 32), ("dec0b", 32, 32))
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet128_l1tsv.npz")


def _bf16_bits(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def pack(params):
    """Weights file arrays: the 3x3 weights of layers 1..13 as bfloat16 bits
    (the precision the kernels compute in), everything else float32."""
    p = np.asarray(params, np.float32)
    w3, f32, o = [], [], 0
    for i, (_, ci, co) in enumerate(LAYERS):
        nw = co * 9 * ci
        (f32 if i == 0 else w3).append(p[o:o + nw])
        f32.append(p[o + nw:o + nw + co])
        o += nw + co
    f32.append(p[o:])
    return {"w3_bf16": _bf16_bits(np.concatenate(w3)), "f32": np.concatenate(f32)}


def unpack(w3_bf16, f32):
    w3 = (np.asarray(w3_bf16, np.uint16).astype(np.uint32) << 16).view(np.float32)
    out, a, b = [], 0, 0
    for i, (_, ci, co) in enumerate(LAYERS):
        nw = co * 9 * ci
        if i == 0:
            out.append(f32[b:b + nw])
            b += nw
        else:
            out.append(w3[a:a + nw])
            a += nw
        out.append(f32[b:b + co])
        b += co
    out.append(f32[b:])
    return np.concatenate(out).astype(np.float32)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise UNetError(f"{LIB_PATH} is missing: build it with python -m smili_amd.build")
    lib = C.CDLL(LIB_PATH)
    lib.smili_unet_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
    lib.smili_unet_destroy.argtypes = [C.c_void_p]
    lib.smili_unet_destroy.restype = None
    lib.smili_unet_nparams.restype = C.c_int64
    lib.smili_unet_set_params.argtypes = [C.c_void_p, _pf]
    lib.smili_unet_get_params.argtypes = [C.c_void_p, _pf]
    lib.smili_unet_forward.argtypes = [C.c_void_p, _pf, C.c_int32, _pf]
    lib.smili_unet_loss_grad.argtypes = [C.c_void_p, _pf, _pf, C.c_int32, _pd, _pf]
    lib.smili_unet_set_data.argtypes = [C.c_void_p, _pf, _pf, C.c_int64]
    lib.smili_unet_train.argtypes = [C.c_void_p, _pl, C.c_int64, _pf, _pf, _pd]
    lib.smili_unet_reset_optimizer.argtypes = [C.c_void_p]
    lib.smili_unet_time_steps.argtypes = [C.c_void_p, C.c_int32, _pf]
    lib.smili_unet_flops.argtypes = [C.c_void_p, _pd]
    lib.smili_unet_time_layer.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _pf, _pd]
    lib.smili_unet_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise UNetError(f"smili_unet error {rc}: {load().smili_unet_last_error().decode()}")


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def nparams():
    return int(load().smili_unet_nparams())


class UNet:
    """Handle of one network on the current HIP device: batches of up to
    ``batch`` images of height x width (multiples of 64 and 128)."""

    def __init__(self, batch=32, height=128, width=128, params=None):
        lib = load()
        h = C.c_void_p()
        _check(lib.smili_unet_create(int(batch), int(height), int(width), C.byref(h)))
        self._h = h
        self.batch, self.height, self.width = int(batch), int(height), int(width)
        if params is not None:
            self.params = params

    def close(self):
        if getattr(self, "_h", None):
            load().smili_unet_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.empty(nparams(), np.float32)
        _check(load().smili_unet_get_params(self._h, out.ctypes.data_as(_pf)))
        return out

    @params.setter
    def params(self, p):
        p = _f32(p, (nparams(),))
        _check(load().smili_unet_set_params(self._h, p.ctypes.data_as(_pf)))

    def forward(self, x):
        """x[n, h, w, 2] -> output[n, h, w] (float32, >= 0)."""
        x = _f32(x)
        if x.ndim != 4 or x.shape[1:] != (self.height, self.width, 2):
            raise ValueError(f"input must be [n, {self.height}, {self.width}, 2], got {x.shape}")
        out = np.empty(x.shape[:3], np.float32)
        _check(load().smili_unet_forward(self._h, x.ctypes.data_as(_pf), x.shape[0], out.ctypes.data_as(_pf)))
        return out

    def loss_grad(self, x, target):
        """Mean squared error of n <= batch images and its gradient."""
        x = _f32(x)
        n = x.shape[0]
        t = _f32(target, (n, self.height, self.width))
        if x.shape != (n, self.height, self.width, 2):
            raise ValueError(f"input must be [n, {self.height}, {self.width}, 2]")
        loss = C.c_double()
        g = np.empty(nparams(), np.float32)
        _check(load().smili_unet_loss_grad(self._h, x.ctypes.data_as(_pf), t.ctypes.data_as(_pf), n,
                                           C.byref(loss), g.ctypes.data_as(_pf)))
        return loss.value, g

    def set_data(self, inputs, targets):
        x = _f32(inputs)
        n = x.shape[0]
        t = _f32(targets, (n, self.height, self.width))
        if x.shape != (n, self.height, self.width, 2):
            raise ValueError(f"inputs must be [n, {self.height}, {self.width}, 2]")
        _check(load().smili_unet_set_data(self._h, x.ctypes.data_as(_pf), t.ctypes.data_as(_pf), n))
        self.ndata = n

    def train(self, index, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        """Adam steps over batches index[step, batch] of the uploaded set;
        an index is 8 * sample + k with k one of the 8 symmetries of the
        square (smili_amd.warmstart.data.dihedral; 0 = identity), applied to
        input and target on the device.  Returns the per-step losses."""
        idx = np.ascontiguousarray(index, dtype=np.int64)
        if idx.ndim != 2 or idx.shape[1] != self.batch:
            raise ValueError(f"index must be [nsteps, {self.batch}]")
        ns = idx.shape[0]
        lr = _f32(np.broadcast_to(np.asarray(lr, np.float32), (ns,)))
        hyper = _f32([beta1, beta2, eps, weight_decay])
        losses = np.empty(ns, np.float64)
        _check(load().smili_unet_train(self._h, idx.ctypes.data_as(_pl), ns, lr.ctypes.data_as(_pf),
                                       hyper.ctypes.data_as(_pf), losses.ctypes.data_as(_pd)))
        return losses

    def reset_optimizer(self):
        _check(load().smili_unet_reset_optimizer(self._h))

    def time_steps(self, reps=10):
        """(ms per training step, ms per forward pass) of a full batch."""
        ms = np.zeros(2, np.float32)
        _check(load().smili_unet_time_steps(self._h, int(reps), ms.ctypes.data_as(_pf)))
        return float(ms[0]), float(ms[1])

    def time_layer(self, layer, pass_=0, reps=20):
        """(mean ms of one launch, FLOP of one launch) of a 3x3 convolution
        kernel: pass 0 forward, 1 input gradient, 2 weight gradient."""
        ms = np.zeros(1, np.float32)
        fl = np.zeros(1, np.float64)
        _check(load().smili_unet_time_layer(self._h, int(layer), int(pass_), int(reps), ms.ctypes.data_as(_pf),
                                            fl.ctypes.data_as(_pd)))
        return float(ms[0]), float(fl[0])

    def flops(self):
        """(3x3-convolution FLOP of a training step, of a forward pass)."""
        f = np.zeros(2, np.float64)
        _check(load().smili_unet_flops(self._h, f.ctypes.data_as(_pd)))
        return float(f[0]), float(f[1])

    def save(self, path, packed=False, **meta):
        arrs = pack(self.params) if packed else {"params": self.params}
        np.savez_compressed(path, height=self.height, width=self.width, **arrs,
                            **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def from_file(cls, path=None, batch=1):
        """(network, metadata) from a weights file (full float32 parameters
        or the packed form); the default is the trained warm-start network."""
        z = np.load(DEFAULT_WEIGHTS if path is None else path)
        net = cls(batch, int(z["height"]), int(z["width"]))
        net.params = z["params"] if "params" in z.files else unpack(z["w3_bf16"], z["f32"])
        skip = ("params", "height", "width", "w3_bf16", "f32")
        return net, {k: z[k] for k in z.files if k not in skip}


def main(argv=None):
    """Times training steps and forward passes of a randomly initialised
    network on random inputs: ``python -m smili_amd.warmstart.unet
    [--batch 32] [--size 128] [--reps 10]``."""
    import argparse
    import json
    import math
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args(argv)
    rng = np.random.default_rng(0)
    parts = []
    for _, ci, co in LAYERS:
        parts += [rng.standard_normal(co * 9 * ci) * math.sqrt(2.0 / (9 * ci)), np.zeros(co)]
    parts += [rng.standard_normal(32) / math.sqrt(32), np.zeros(1)]
    with UNet(a.batch, a.size, a.size, params=np.concatenate(parts)) as net:
        x = rng.standard_normal((a.batch, a.size, a.size, 2)).astype(np.float32)
        net.set_data(x, np.abs(x[..., 0]))
        net.time_steps(2)
        step_ms, fwd_ms = net.time_steps(a.reps)
        f_step, f_fwd = net.flops()
    print(json.dumps({"batch": a.batch, "size": a.size, "train_step_ms": round(step_ms, 4),
                      "train_TFLOPs": round(f_step / step_ms / 1e9, 1), "forward_ms": round(fwd_ms, 4),
                      "forward_TFLOPs": round(f_fwd / fwd_ms / 1e9, 1)}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
