"""Warm-start U-Net on the GPU (libsmili_unet.so through its C ABI) against
the CPU oracle (oracle/unet_oracle.py, which emulates the same bfloat16
rounding points and sums in float64).

Both sides round activations to bfloat16 at the same places; they differ
only where a float32 sum on the GPU and the float64 sum of the oracle fall
on different sides of a bfloat16 rounding boundary (one bfloat16 ulp,
2^-8 relative, at a small fraction of the elements), so the bars are:
output rel-l2 <= 3e-2 and below the oracle's own bfloat16-vs-exact
difference (about 1% of the activations of a layer flip by one bfloat16
ulp, and the head's 32-term sum cancels), loss rel <= 5e-3, gradient rel-l2 <= 1e-2 over all parameters
and <= 3e-2 per layer.

A network with sparse ternary integer weights, small integer inputs and
integer targets has every product and every short sum exact in float32 and
float64 (each output of the forward pass and of the input gradients has a
few nonzero terms), so every bfloat16 rounding sees the same value on both
sides: there the GPU forward pass and loss must equal the oracle's bit for
bit, and the weight gradients (sums over all pixels, where float32 order
matters) agree to 1e-5 per layer.  This pins every index of every kernel."""
import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _inputs(rng, n, h, w):
    x = np.empty((n, h, w, 2), np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    for b in range(n):
        cy, cx = rng.uniform(0.3, 0.7) * h, rng.uniform(0.3, 0.7) * w
        x[b, :, :, 0] = np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * (0.1 * h) ** 2)) + 0.1 * rng.standard_normal((h, w))
        x[b, :, :, 1] = np.exp(-((yy - h / 2) ** 2 + (xx - w / 2) ** 2) / (2 * 3.0 ** 2))
    return x


@pytest.fixture(scope="module")
def net_case(gpu_lib):
    from smili_amd.warmstart import unet
    rng = np.random.default_rng(11)
    theta = U.init_params(rng)
    theta[-1] = 0.2
    # nonzero biases exercise the bias paths
    parts = U.unpack(theta)
    for name, (W, b) in parts.items():
        b[...] = 0.05 * rng.standard_normal(b.shape)
    net = unet.UNet(batch=2, height=64, width=128, params=theta)
    yield net, theta, rng
    net.close()


def test_forward_matches_oracle(net_case):
    net, theta, rng = net_case
    x = _inputs(rng, 2, 64, 128)
    out = net.forward(x)
    ref = U.forward(theta, x)
    exact = U.forward(theta, x, bf16_mode=False)
    print("forward rel-l2", rel_l2(out, ref), "oracle bf16 vs exact", rel_l2(ref, exact))
    assert out.shape == (2, 64, 128) and np.all(out >= 0)
    assert rel_l2(out, ref) <= min(3e-2, rel_l2(ref, exact))


def _integer_params(rng):
    parts = []
    for _, ci, co in U.layer_specs():
        W = rng.choice([-1.0, 0.0, 1.0], size=co * 9 * ci, p=[1.0 / (9 * ci), 1 - 2.0 / (9 * ci), 1.0 / (9 * ci)])
        parts += [W, rng.integers(0, 2, co).astype(np.float64)]
    parts += [rng.choice([-1.0, 0.0, 1.0], size=32), np.array([1.0])]
    return np.concatenate(parts).astype(np.float32)


def test_integer_network_is_exact(gpu_lib):
    from smili_amd.warmstart import unet
    rng = np.random.default_rng(12)
    theta = _integer_params(rng)
    x = np.stack([rng.integers(0, 4, (2, 64, 128)), rng.integers(0, 2, (2, 64, 128))], axis=-1).astype(np.float32)
    t = rng.integers(0, 3, (2, 64, 128)).astype(np.float32)
    with unet.UNet(batch=2, height=64, width=128, params=theta) as net:
        out = net.forward(x)
        L, g = net.loss_grad(x, t)
    ref = U.forward(theta, x)
    Lr, gr = U.loss_grad(theta, x, t)
    print("integer net: output range", out.min(), out.max(), "nonzero grads", np.count_nonzero(gr),
          "bit-equal grads", np.array_equal(g, gr.astype(np.float32)))
    assert np.count_nonzero(gr) > 1000 and out.max() > 0
    assert np.array_equal(out, ref.astype(np.float32))
    assert abs(L - Lr) <= 1e-12 * Lr
    o = 0
    for name, ci, co in U.layer_specs():
        n = co * 9 * ci + co
        assert rel_l2(g[o:o + n], gr[o:o + n]) <= 1e-5, name
        o += n
    assert rel_l2(g[o:], gr[o:]) <= 1e-5


def test_forward_batches_a_remainder(net_case):
    net, theta, rng = net_case
    x = _inputs(rng, 3, 64, 128)
    out = net.forward(x)
    assert rel_l2(out[2:], U.forward(theta, x[2:])) <= 3e-2
    assert np.array_equal(out[:2], net.forward(x[:2]))


def test_loss_grad_matches_oracle(net_case):
    """The target is a smooth blob independent of the output, so the
    residual has a coherent sign and every parameter gradient is a coherent
    sum over pixels (a residual of random sign would make the gradients
    random-walk sums, where the small forward differences weigh sqrt(N)
    times more)."""
    net, theta, rng = net_case
    x = _inputs(rng, 2, 64, 128)
    yy, xx = np.mgrid[0:64, 0:128]
    blob = 2.0 * np.exp(-((yy - 30) ** 2 + (xx - 70) ** 2) / (2 * 12.0 ** 2))
    t = np.stack([blob, blob[::-1]]).astype(np.float32)
    L, g = net.loss_grad(x, t)
    Lr, gr = U.loss_grad(theta, x, t)
    print("loss", L, Lr, "grad rel-l2", rel_l2(g, gr))
    errs, o = {}, 0
    for name, ci, co in U.layer_specs():
        nw = co * 9 * ci
        errs[name] = (rel_l2(g[o:o + nw], gr[o:o + nw]), rel_l2(g[o + nw:o + nw + co], gr[o + nw:o + nw + co]))
        print(f"  {name}: W {errs[name][0]:.2e}  b {errs[name][1]:.2e}")
        o += nw + co
    errs["head"] = (rel_l2(g[o:], gr[o:]), 0.0)
    print(f"  head: {errs['head'][0]:.2e}")
    assert abs(L - Lr) <= 5e-3 * Lr
    assert rel_l2(g, gr) <= 1e-2
    for name, (ew, eb) in errs.items():
        assert ew <= 3e-2 and eb <= 3e-2, name


def test_adam_step_and_training_reduces_loss(net_case):
    net, theta, rng = net_case
    from smili_amd.warmstart import unet
    x = _inputs(rng, 4, 64, 128)
    t = np.zeros((4, 64, 128), np.float32)
    t[:, 24:40, 48:80] = 1.0
    with unet.UNet(batch=2, height=64, width=128, params=theta) as n2:
        n2.set_data(x, t)
        L0, g0 = n2.loss_grad(x[:2], t[:2])
        lr = 1e-3
        losses = n2.train(8 * np.array([[0, 1]]), lr)
        assert abs(losses[0] - L0) <= 1e-4 * L0
        p1 = n2.params
        ref, _, _ = U.adam_step(theta.astype(np.float64), 0.0, 0.0, g0.astype(np.float64), 1, lr)
        # first Adam step moves every parameter by lr * g / (|g| + eps)
        bad = np.abs(p1 - ref) > 1e-3 * lr + 1e-6 * np.abs(ref)
        assert bad.mean() <= 1e-4, bad.mean()
        for bad in (dict(beta1=1.0), dict(beta2=-0.1), dict(eps=0.0), dict(weight_decay=-1.0)):
            with pytest.raises(unet.UNetError, match="hyper"):
                n2.train(8 * np.array([[0, 1]]), lr, **bad)
        with pytest.raises(unet.UNetError, match="index out of range"):
            n2.train(np.array([[0, 8 * 4]]), lr)
        assert np.array_equal(n2.params, p1)
        idx = 8 * np.tile(np.array([[0, 1], [2, 3]]), (30, 1))
        losses = n2.train(idx, 2e-4)
        print("train losses", losses[:2], losses[-2:])
        assert np.all(np.isfinite(losses)) and losses[-2:].mean() < 0.5 * losses[:2].mean()


def test_symmetry_index_matches_host_transform(gpu_lib):
    """Training index 8 s + k applies the k-th symmetry of the square on
    the device; its loss equals the loss of the host-transformed sample
    (smili_amd.warmstart.data.dihedral)."""
    from smili_amd.warmstart import unet
    from smili_amd.warmstart.data import dihedral
    rng = np.random.default_rng(13)
    theta = U.init_params(rng)
    x = rng.standard_normal((2, 128, 128, 2)).astype(np.float32)
    t = np.abs(rng.standard_normal((2, 128, 128))).astype(np.float32)
    with unet.UNet(batch=1, height=128, width=128, params=theta) as net:
        net.set_data(x, t)
        for k in range(8):
            L = net.train(np.array([[8 * 1 + k]]), 0.0)[0]
            xr = np.ascontiguousarray(dihedral(x[1], k))[None]
            tr = np.ascontiguousarray(dihedral(t[1], k))[None]
            Lr, _ = net.loss_grad(xr, tr)
            assert abs(L - Lr) <= 1e-9 * Lr, (k, L, Lr)
        assert np.array_equal(net.params, theta)


def test_time_steps_leave_the_training_state(net_case):
    """Timing steps in the middle of training leave the parameters, the Adam
    moments and the step count as they were: training on afterwards gives
    the same parameters as training without the timing (to the float
    reductions' rounding)."""
    net, theta, rng = net_case
    from smili_amd.warmstart import unet
    x = _inputs(rng, 4, 64, 128)
    t = np.zeros((4, 64, 128), np.float32)
    t[:, 24:40, 48:80] = 1.0
    idx = 8 * np.array([[0, 1], [2, 3], [1, 2]])
    out = []
    for timed in (False, True):
        with unet.UNet(batch=2, height=64, width=128, params=theta) as n2:
            n2.set_data(x, t)
            n2.train(idx[:2], 1e-3)
            if timed:
                p_before = n2.params
                n2.time_steps(3)
                assert np.array_equal(n2.params, p_before)
            n2.train(idx[2:], 1e-3)
            out.append(n2.params)
    rel = np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0])
    print("params after timing vs without", rel)
    assert rel <= 1e-6
