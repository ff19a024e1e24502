"""Pins of the warm-start U-Net oracle (oracle/unet_oracle.py) against
what fixes it independently: torch's bfloat16 cast (round to nearest even),
brute-force loops and torch.nn.functional.conv2d for the 3x3 convolution,
adjoint identities for its two backward products, central finite
differences of the whole loss in the unrounded float64 mode, and
torch.optim.AdamW for the optimiser step."""
import numpy as np
import pytest
import torch

from oracle import unet_oracle as U


def test_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5e-3, 3.14159, 1e30, -7.77e-20, 0.0], np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(U.bf16(x), ref)
    assert U.bf16(x)[1] == 1.0 and U.bf16(x)[2] == 1 + 2 ** -6      # ties to even
    rng = np.random.default_rng(0)
    y = (rng.standard_normal(100000) * 10.0 ** rng.uniform(-20, 20, 100000)).astype(np.float32)
    assert np.array_equal(U.bf16(y), torch.from_numpy(y).to(torch.bfloat16).float().numpy())


def test_conv3x3_brute_force():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2, 4, 5, 3))
    W = rng.standard_normal((2, 9, 3))
    b = rng.standard_normal(2)
    Y = U.conv3x3(X, W, b)
    ref = np.zeros_like(Y)
    for n in range(2):
        for y in range(4):
            for x in range(5):
                for co in range(2):
                    s = b[co]
                    for dy in range(3):
                        for dx in range(3):
                            yy, xx = y + dy - 1, x + dx - 1
                            if 0 <= yy < 4 and 0 <= xx < 5:
                                s += W[co, 3 * dy + dx] @ X[n, yy, xx]
                    ref[n, y, x, co] = s
    assert np.abs(Y - ref).max() <= 1e-12


def test_conv3x3_vs_torch_and_adjoints():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((2, 8, 16, 6))
    W = rng.standard_normal((5, 9, 6))
    b = rng.standard_normal(5)
    Y = U.conv3x3(X, W, b)
    Wt = torch.from_numpy(W.reshape(5, 3, 3, 6).transpose(0, 3, 1, 2).copy())
    ref = torch.nn.functional.conv2d(torch.from_numpy(X.transpose(0, 3, 1, 2).copy()), Wt,
                                     torch.from_numpy(b), padding=1).numpy().transpose(0, 2, 3, 1)
    assert np.abs(Y - ref).max() <= 1e-12
    # <conv(X), D> = <X, conv_dx(D)> = <W, conv_dw(D, X)>  (no bias)
    D = rng.standard_normal(Y.shape)
    lhs = np.sum(U.conv3x3(X, W, np.zeros(5)) * D)
    assert abs(lhs - np.sum(X * U.conv3x3_dx(D, W))) <= 1e-10 * abs(lhs)
    assert abs(lhs - np.sum(W * U.conv3x3_dw(D, X))) <= 1e-10 * abs(lhs)


def test_pool_up_adjoints():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1, 8, 8, 3))
    Yp = rng.standard_normal((1, 4, 4, 3))
    assert np.isclose(np.sum(U.pool(X) * Yp), np.sum(X * U.pool_dx(Yp)))
    Xs = rng.standard_normal((1, 4, 4, 3))
    Yu = rng.standard_normal((1, 8, 8, 3))
    assert np.isclose(np.sum(U.up(Xs) * Yu), np.sum(Xs * U.up_dx(Yu)))


@pytest.mark.parametrize("base", [4, 32])
def test_loss_grad_finite_differences(base):
    rng = np.random.default_rng(4 + base)
    theta = U.init_params(rng, base).astype(np.float64)
    theta[-1] = 0.3
    x = rng.standard_normal((2, 16, 16, 2))
    t = np.abs(rng.standard_normal((2, 16, 16)))
    L, g = U.loss_grad(theta, x, t, base, bf16_mode=False)
    assert g.size == U.nparams(base) == theta.size
    for k in range(3):
        d = rng.standard_normal(theta.size)
        eps = 1e-6 / np.linalg.norm(d) * np.linalg.norm(theta)
        Lp, _ = U.loss_grad(theta + eps * d, x, t, base, bf16_mode=False)
        Lm, _ = U.loss_grad(theta - eps * d, x, t, base, bf16_mode=False)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-6 * abs(g @ d) + 1e-12, (k, fd, g @ d)
    # single coordinates in the first, a middle and the head layer
    for i in (5, U.nparams(base) // 2, U.nparams(base) - 2):
        e = np.zeros_like(theta)
        e[i] = 1e-5
        fd = (U.loss_grad(theta + e, x, t, base, False)[0] - U.loss_grad(theta - e, x, t, base, False)[0]) / 2e-5
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g[i]), 1e-3 * np.abs(g).max()), (i, fd, g[i])


def test_bf16_mode_close_to_exact():
    """The rounded network stays within bfloat16 noise of the exact one."""
    rng = np.random.default_rng(6)
    theta = U.init_params(rng)
    x = rng.standard_normal((1, 16, 16, 2)).astype(np.float32)
    a = U.forward(theta, x, bf16_mode=True)
    b = U.forward(theta, x, bf16_mode=False)
    assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b)


def test_adam_matches_torch_adamw():
    rng = np.random.default_rng(7)
    p0 = rng.standard_normal(50)
    th = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.AdamW([th], lr=1e-2, betas=(0.9, 0.99), eps=1e-8, weight_decay=1e-3)
    theta, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.standard_normal(50)
        th.grad = torch.from_numpy(g.copy())
        opt.step()
        theta, m, v = U.adam_step(theta, m, v, g, step, 1e-2, 0.9, 0.99, 1e-8, 1e-3)
    assert np.abs(theta - th.detach().numpy()).max() <= 1e-12
