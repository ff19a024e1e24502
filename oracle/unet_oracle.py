"""CPU oracle of the warm-start U-Net (test infrastructure only: tests,
``__graft_entry__.smoke()`` and the bench's CPU leg use it as the checker;
the product path never imports it).

The network has no counterpart in the reference (SMILI images by L-BFGS-B
only); it is this project's warm-start proposal, specified here in full so
the GPU kernels have an independent restatement to be checked against:

* input x[B, H, W, 2] float32 (channel 0: dirty image / its peak, channel 1:
  dirty beam / its peak), H % 64 == 0, W % 128 == 0;
* encoder level l = 0..3 at H/2^l x W/2^l with C_l = 32, 64, 128, 256
  channels: two 3x3 convolutions (zero padding, + bias, ReLU); between
  levels a 2x2 average pool;
* decoder level l = 2, 1, 0: the first 3x3 convolution reads the channel
  concatenation [nearest x2 upsample of the level below (C_{l+1}), skip of
  encoder level l (C_l)], the second reads C_l; both ReLU;
* head: z = sum_c w_c d0[c] + b; the image is max(z, 0) (non-negative);
* loss L = (1/B) sum_b (1/(H W)) sum_px (z - target)^2 on the linear head
  (targets are non-negative, so clamping can only lower the error; a
  saturating output such as softplus stops learning where most target
  pixels are zero).

Parameters (flat float32 vector, in order): for each 3x3 convolution
W[co][tap][ci] (tap = 3 dy + dx) then bias[co] -- conv_in (2 -> 32), enc0b,
enc1a, enc1b, enc2a, enc2b, enc3a, enc3b, dec2a, dec2b, dec1a, dec1b, dec0a,
dec0b -- then the head's w[32] and b[1].

Rounding points of the GPU path (``bf16=True``): 3x3 weights except
conv_in's are rounded to bfloat16 (round to nearest even); every ReLU
output and every pooled tensor is stored in bfloat16; conv_in, the head and
all gradient tensors of activations stay float32, except that the
masked output gradient dP = dA * (Y > 0) entering a 3x3 convolution's
backward is rounded to bfloat16.  Sums here run in float64.
With ``bf16=False`` nothing is rounded (the float64 mode the finite
difference checks use).
"""
from __future__ import annotations

import numpy as np

BASE = 32


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as
    float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def layer_specs(base=BASE):
    """[(name, cin, cout)] of the 3x3 convolutions, in parameter order."""
    c = [base, 2 * base, 4 * base, 8 * base]
    return [("conv_in", 2, c[0]), ("enc0b", c[0], c[0]),
            ("enc1a", c[0], c[1]), ("enc1b", c[1], c[1]),
            ("enc2a", c[1], c[2]), ("enc2b", c[2], c[2]),
            ("enc3a", c[2], c[3]), ("enc3b", c[3], c[3]),
            ("dec2a", c[3] + c[2], c[2]), ("dec2b", c[2], c[2]),
            ("dec1a", c[2] + c[1], c[1]), ("dec1b", c[1], c[1]),
            ("dec0a", c[1] + c[0], c[0]), ("dec0b", c[0], c[0])]


def nparams(base=BASE):
    return sum(co * 9 * ci + co for _, ci, co in layer_specs(base)) + base + 1


def unpack(theta, base=BASE):
    """Flat parameters -> {name: (W[co, 9, ci], b[co])}, head (w, b)."""
    theta = np.asarray(theta)
    out, o = {}, 0
    for name, ci, co in layer_specs(base):
        W = theta[o:o + co * 9 * ci].reshape(co, 9, ci)
        o += co * 9 * ci
        out[name] = (W, theta[o:o + co])
        o += co
    out["head"] = (theta[o:o + base], theta[o + base:o + base + 1])
    assert o + base + 1 == theta.size
    return out


def init_params(rng, base=BASE):
    """He-normal 3x3 weights, zero biases, small head (float32)."""
    parts = []
    for _, ci, co in layer_specs(base):
        parts.append(rng.standard_normal(co * 9 * ci) * np.sqrt(2.0 / (9 * ci)))
        parts.append(np.zeros(co))
    parts.append(rng.standard_normal(base) * np.sqrt(1.0 / base))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


# ----------------------------------------------------------------------------
# building blocks (NHWC)
# ----------------------------------------------------------------------------
def _shifts(H, W):
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        yield t, dy, dx


def conv3x3(X, Wt, b):
    """Y[b, y, x, co] = sum_{tap, ci} W[co, tap, ci] X[b, y+dy, x+dx, ci] + b[co]
    with zero padding (float64)."""
    X = np.asarray(X, np.float64)
    B, H, W, C = X.shape
    P = np.zeros((B, H + 2, W + 2, C))
    P[:, 1:-1, 1:-1] = X
    Y = np.zeros((B, H, W, Wt.shape[0]))
    for t, dy, dx in _shifts(H, W):
        Y += np.tensordot(P[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W], np.asarray(Wt[:, t, :], np.float64).T, 1)
    return Y + np.asarray(b, np.float64)


def conv3x3_dx(dP, Wt):
    """Adjoint of conv3x3 in its input: dX[q] = sum_tap W[tap]^T dP[q - off]."""
    dP = np.asarray(dP, np.float64)
    B, H, W, _ = dP.shape
    P = np.zeros((B, H + 2, W + 2, dP.shape[3]))
    P[:, 1:-1, 1:-1] = dP
    dX = np.zeros((B, H, W, Wt.shape[2]))
    for t, dy, dx in _shifts(H, W):
        dX += np.tensordot(P[:, 1 - dy:1 - dy + H, 1 - dx:1 - dx + W], np.asarray(Wt[:, t, :], np.float64), 1)
    return dX


def conv3x3_dw(dP, X):
    """dW[co, tap, ci] = sum_{b, p} dP[b, p, co] X[b, p + off, ci]."""
    X = np.asarray(X, np.float64)
    dP = np.asarray(dP, np.float64)
    B, H, W, C = X.shape
    P = np.zeros((B, H + 2, W + 2, C))
    P[:, 1:-1, 1:-1] = X
    dW = np.zeros((dP.shape[3], 9, C))
    for t, dy, dx in _shifts(H, W):
        dW[:, t, :] = np.tensordot(dP, P[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W], ([0, 1, 2], [0, 1, 2]))
    return dW


def pool(X):
    B, H, W, C = X.shape
    return np.asarray(X, np.float64).reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_dx(dY):
    return np.repeat(np.repeat(dY, 2, axis=1), 2, axis=2) / 4.0


def up(X):
    return np.repeat(np.repeat(X, 2, axis=1), 2, axis=2)


def up_dx(dY):
    B, H, W, C = dY.shape
    return dY.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


# ----------------------------------------------------------------------------
# network
# ----------------------------------------------------------------------------
def forward(theta, x, base=BASE, bf16_mode=True, keep=False):
    """Network image max(z, 0) [B, H, W] (and the stored activations, with
    the linear head z, if keep)."""
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    P = unpack(theta, base)

    def wq(name):
        W, b = P[name]
        if name == "conv_in":
            return np.asarray(W, np.float64), b
        return r(W), b

    A = {}
    x = np.asarray(x, np.float32)

    def cr(name, X):
        W, b = wq(name)
        return r(np.maximum(conv3x3(X, W, b), 0.0))

    A["x"] = x
    A["e0a"] = cr("conv_in", x)
    A["e0"] = cr("enc0b", A["e0a"])
    A["p0"] = r(pool(A["e0"]))
    A["e1a"] = cr("enc1a", A["p0"])
    A["e1"] = cr("enc1b", A["e1a"])
    A["p1"] = r(pool(A["e1"]))
    A["e2a"] = cr("enc2a", A["p1"])
    A["e2"] = cr("enc2b", A["e2a"])
    A["p2"] = r(pool(A["e2"]))
    A["e3a"] = cr("enc3a", A["p2"])
    A["e3"] = cr("enc3b", A["e3a"])
    A["c2"] = np.concatenate([up(A["e3"]), A["e2"]], axis=3)
    A["d2a"] = cr("dec2a", A["c2"])
    A["d2"] = cr("dec2b", A["d2a"])
    A["c1"] = np.concatenate([up(A["d2"]), A["e1"]], axis=3)
    A["d1a"] = cr("dec1a", A["c1"])
    A["d1"] = cr("dec1b", A["d1a"])
    A["c0"] = np.concatenate([up(A["d1"]), A["e0"]], axis=3)
    A["d0a"] = cr("dec0a", A["c0"])
    A["d0"] = cr("dec0b", A["d0a"])
    hw, hb = P["head"]
    A["z"] = np.tensordot(np.asarray(A["d0"], np.float64), np.asarray(hw, np.float64), 1) + float(hb[0])
    out = np.maximum(A["z"], 0.0)
    return (out, A) if keep else out


def loss_grad(theta, x, target, base=BASE, bf16_mode=True):
    """(L, dL/dtheta float64 flat) of the mean squared error."""
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    _, A = forward(theta, x, base, bf16_mode, keep=True)
    B, H, W = A["z"].shape
    diff = A["z"] - np.asarray(target, np.float64)
    L = float(np.sum(diff ** 2) / (B * H * W))
    P = unpack(theta, base)
    G = {}
    dz = 2.0 * diff / (B * H * W)
    hw, _ = P["head"]
    G["head"] = (np.tensordot(dz, np.asarray(A["d0"], np.float64), ([0, 1, 2], [0, 1, 2])),
                 np.array([dz.sum()]))
    dA = dz[..., None] * np.asarray(hw, np.float64)

    def back(name, Y, X, dA):
        """Through ReLU and one 3x3 convolution: returns dX (float64)."""
        W, _ = P[name]
        Wq = np.asarray(W, np.float64) if name == "conv_in" else r(W)
        g = dA * (Y > 0)
        G_b = g.sum(axis=(0, 1, 2))
        dP = r(g)
        G[name] = (conv3x3_dw(dP, X), G_b)
        return conv3x3_dx(dP, Wq)

    dd0a = back("dec0b", A["d0"], A["d0a"], dA)
    dc0 = back("dec0a", A["d0a"], A["c0"], dd0a)
    C1 = A["d1"].shape[3]
    dd1 = up_dx(dc0[..., :C1])
    de0 = dc0[..., C1:].copy()
    dd1a = back("dec1b", A["d1"], A["d1a"], dd1)
    dc1 = back("dec1a", A["d1a"], A["c1"], dd1a)
    C2 = A["d2"].shape[3]
    dd2 = up_dx(dc1[..., :C2])
    de1 = dc1[..., C2:].copy()
    dd2a = back("dec2b", A["d2"], A["d2a"], dd2)
    dc2 = back("dec2a", A["d2a"], A["c2"], dd2a)
    C3 = A["e3"].shape[3]
    de3 = up_dx(dc2[..., :C3])
    de2 = dc2[..., C3:].copy()
    de3a = back("enc3b", A["e3"], A["e3a"], de3)
    dp2 = back("enc3a", A["e3a"], A["p2"], de3a)
    de2 += pool_dx(dp2)
    de2a = back("enc2b", A["e2"], A["e2a"], de2)
    dp1 = back("enc2a", A["e2a"], A["p1"], de2a)
    de1 += pool_dx(dp1)
    de1a = back("enc1b", A["e1"], A["e1a"], de1)
    dp0 = back("enc1a", A["e1a"], A["p0"], de1a)
    de0 += pool_dx(dp0)
    de0a = back("enc0b", A["e0"], A["e0a"], de0)
    # conv_in: its input needs no gradient
    W, _ = P["conv_in"]
    g = de0a * (A["e0a"] > 0)
    G["conv_in"] = (conv3x3_dw(r(g), A["x"]), g.sum(axis=(0, 1, 2)))
    parts = []
    for name, _, _ in layer_specs(base):
        parts += [G[name][0].reshape(-1), G[name][1].reshape(-1)]
    parts += [G["head"][0].reshape(-1), G["head"][1].reshape(-1)]
    return L, np.concatenate(parts)


def adam_step(theta, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0):
    """One Adam update (Kingma & Ba 2015, with bias correction; decoupled
    weight decay wd, Loshchilov & Hutter 2019) in float64; ``step`` counts
    from 1.  Returns (theta, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    theta = theta - lr * (mh / (np.sqrt(vh) + eps) + wd * theta)
    return theta, m, v
