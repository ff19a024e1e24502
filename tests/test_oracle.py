"""Pins of the FP64 oracle (CPU).  Each test would fail for a plausible
mistake in the oracle: a wrong sign or mode offset in the transforms, a
dropped term or wrong factor in a chi-square gradient, a transposed stencil
or a missing quirk in the regularizers."""
import math

import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import random_problem, rel_l2


# ----------------------------------------------------------------------------
# transforms
# ----------------------------------------------------------------------------
def test_dft_matches_brute_force_loops():
    rng = np.random.default_rng(0)
    Nx, Ny = 7, 6
    I = rng.standard_normal((Ny, Nx))
    u = rng.uniform(-3, 3, 5)
    v = rng.uniform(-3, 3, 5)
    V = np.zeros(5, complex)
    for j in range(5):
        for iy in range(1, Ny + 1):
            for ix in range(1, Nx + 1):
                # wrapfinufft.f90:19,31: mode k = ix - 1 - N/2 (integer division)
                V[j] += I[iy - 1, ix - 1] * np.exp(1j * ((ix - 1 - Nx // 2) * u[j] + (iy - 1 - Ny // 2) * v[j]))
    assert np.abs(O.dft_fwd(I, u, v) - V).max() < 1e-12
    c = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    F = np.zeros((Ny, Nx), complex)
    for iy in range(1, Ny + 1):
        for ix in range(1, Nx + 1):
            F[iy - 1, ix - 1] = sum(c[j] * np.exp(1j * ((ix - 1 - Nx // 2) * u[j] + (iy - 1 - Ny // 2) * v[j]))
                                    for j in range(5))
    assert np.abs(O.dft_adj(c, u, v, Nx, Ny) - F).max() < 1e-12


def test_dft_with_phase_shift_equals_reference_dft_matrix():
    """The reference's direct-DFT matrix (dftlib.f90:24-53) is
    exp(+i 2 pi (u x + v y)) with sky positions x = dx (ix - Nxref)
    (imdata.py:743-744).  The NUFFT model with tracking centre at pixel
    N/2+1 equals it times exp(+i(u' dix + v' diy)), dix = Nxref - Nx/2 - 1
    (phashift_r2c, fftlib.f90:223-244)."""
    rng = np.random.default_rng(1)
    Nx, Ny = 12, 10
    Nxref, Nyref = 5.5, 7.0
    dx, dy = -2e-11, 3e-11
    I = rng.standard_normal((Ny, Nx))
    u = rng.uniform(-5e9, 5e9, 40)
    v = rng.uniform(-5e9, 5e9, 40)
    xs = dx * (np.arange(1, Nx + 1) - Nxref)
    ys = dy * (np.arange(1, Ny + 1) - Nyref)
    X, Y = np.meshgrid(xs, ys)
    Vref = np.array([np.sum(I * np.exp(1j * 2 * np.pi * (uu * X + vv * Y))) for uu, vv in zip(u, v)])
    up, vp = 2 * np.pi * dx * u, 2 * np.pi * dy * v
    Vn = O.dft_fwd(I, up, vp)
    idx = np.arange(1, 41, dtype=np.int32)
    shifted = O.phashift_r2c(Vref, idx, up, vp, Nxref, Nyref, Nx, Ny)
    assert np.abs(shifted - Vn).max() < 1e-10 * np.abs(Vn).max()


def test_gaussian_image_matches_closed_form():
    from smili_amd import synth
    dx = dy = 2 * synth.UAS
    sky = [synth.Gaussian(1.0, 30 * synth.UAS, 10 * synth.UAS, -6 * synth.UAS)]
    img = synth.sky_image(sky, 64, 64, -dx, dy)
    rng = np.random.default_rng(2)
    u = rng.uniform(-1e9, 1e9, 30)
    v = rng.uniform(-1e9, 1e9, 30)
    Vt = synth.sky_vis(sky, u, v)
    up, vp = 2 * np.pi * -dx * u, 2 * np.pi * dy * v
    Vn = O.dft_fwd(img, up, vp)
    # shift the analytic visibilities to the NUFFT tracking centre (pixel 33)
    Vs = O.phashift_r2c(Vt, np.arange(1, 31), up, vp, 32.5, 32.5, 64, 64)
    assert np.abs(Vs - Vn).max() < 2e-3 * np.abs(Vt).max()


def test_adjoint_identity():
    rng = np.random.default_rng(3)
    Nx, Ny = 9, 8
    x = rng.standard_normal((Ny, Nx))
    u = rng.uniform(-3, 3, 20)
    v = rng.uniform(-3, 3, 20)
    y = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    lhs = np.vdot(y, O.dft_fwd(x, u, v)).real
    rhs = np.sum(x * O.dft_adj(np.conj(y), u, v, Nx, Ny).real)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize("shape", [(64, 48), (33, 17), (40, 40)])
def test_es_nufft_within_eps_of_dft(shape):
    rng = np.random.default_rng(4)
    Nx, Ny = shape
    I = rng.standard_normal((Ny, Nx))
    u = rng.uniform(-np.pi, np.pi, 2000)
    v = rng.uniform(-np.pi, np.pi, 2000)
    assert rel_l2(O.es_nufft_type2(I, u, v), O.dft_fwd(I, u, v)) < 2e-6
    c = rng.standard_normal(2000) + 1j * rng.standard_normal(2000)
    assert rel_l2(O.es_nufft_type1(c, u, v, Nx, Ny), O.dft_adj(c, u, v, Nx, Ny)) < 2e-6


def test_es_correction_matches_direct_kernel_transform():
    """c(k) is 1/(kernel Fourier transform): check against brute-force
    trapezoid integration of phi."""
    n, nf = 16, 32
    c = O.es_correction(n, nf)
    z = np.linspace(-1, 1, 200001)
    phi = O.es_kernel(z)
    for i, k in enumerate(O.mode_indices(n)):
        ft = (O.ES_WIDTH / 2) * np.trapezoid(phi * np.cos(np.pi * k * O.ES_WIDTH * z / nf), z)
        assert abs(1 / ft - c[i]) < 1e-8 * abs(c[i])


# ----------------------------------------------------------------------------
# chi-square terms
# ----------------------------------------------------------------------------
def _fd_grad_check(p, x, n_dir=3, h=1e-6, rtol=1e-6, which="chisq"):
    rng = np.random.default_rng(5)
    f = (lambda z: O.calc_chisq(z, p)[0]) if which == "chisq" else (lambda z: O.calc_cost_reg(z, p)[0])
    g = O.calc_chisq(x, p)[1] if which == "chisq" else O.calc_cost_reg(x, p)[1]
    for _ in range(n_dir):
        d = rng.standard_normal(x.size)
        fd = (f(x + h * d) - f(x - h * d)) / (2 * h)
        an = g @ d
        assert abs(fd - an) <= rtol * max(abs(an), 1e-12) + 1e-9, (fd, an)


@pytest.mark.parametrize("only", ["fcv", "amp", "cp", "ca"])
def test_chisq_gradient_finite_differences(only):
    rng = np.random.default_rng({"fcv": 10, "amp": 11, "cp": 12, "ca": 13}[only])
    n = dict(nfcv=0, namp=0, ncp=0, nca=0)
    n["n" + only] = 40
    p, x = random_problem(rng, 8, 7, M=30, regs=(), **n)
    _fd_grad_check(p, x, rtol=1e-5)


@pytest.mark.parametrize("only", ["fcv", "amp", "cp", "ca"])
def test_chisq_zero_at_noise_free_data(only):
    rng = np.random.default_rng(20)
    n = dict(nfcv=0, namp=0, ncp=0, nca=0)
    n["n" + only] = 50
    p, x = random_problem(rng, 8, 8, M=40, regs=(), noise=0.0, **n)
    chisq, grad, *_ = O.calc_chisq(p.extra["truth"], p)
    assert chisq < 1e-20
    assert np.abs(grad).max() < 1e-9


def test_closure_terms_invariant_under_translation():
    """Closure phases and log closure amplitudes of closed triangles /
    quadrangles do not depend on the image position."""
    rng = np.random.default_rng(6)
    Nx = Ny = 16
    I = np.zeros((Ny, Nx))
    I[5:9, 4:8] = rng.uniform(0.5, 1, (4, 4))
    I2 = np.roll(np.roll(I, 3, axis=1), 2, axis=0)
    # closed triangles: b3 = -(b1 + b2), quadrangles from 4 station positions
    st = rng.uniform(-0.15, 0.15, (5, 2))
    tri = [(0, 1), (1, 2), (2, 0)]
    quad = [(0, 1), (2, 3), (0, 2), (1, 3)]
    uvl = [st[b] - st[a] for a, b in tri + quad]
    u = np.array([q[0] for q in uvl])
    v = np.array([q[1] for q in uvl])
    sgn = np.where(v < 0, -1, 1)
    u, v = u * sgn, v * sgn
    idx = (np.arange(7) + 1) * sgn
    kw = dict(Nx=Nx, Ny=Ny, xidx=np.tile(np.arange(1, Nx + 1), Ny), yidx=np.repeat(np.arange(1, Ny + 1), Nx),
              u=u, v=v, uvidxcp=idx[:3].reshape(3, 1).astype(np.int32), CP=np.array([0.3]),
              Varcp=np.array([0.1]), wcp=0.5, uvidxca=idx[3:].reshape(4, 1).astype(np.int32),
              CA=np.array([0.2]), Varca=np.array([0.1]), wca=0.5)
    p = O.Problem(**kw)
    c1 = O.calc_chisq(I.reshape(-1), p)[0]
    c2 = O.calc_chisq(I2.reshape(-1), p)[0]
    assert abs(c1 - c2) < 1e-10 * abs(c1)


def test_fcv_conjugate_index():
    """A negative index means the datum sits at (-u, -v): the model is
    conj(V) (fftlib.f90:424-428)."""
    rng = np.random.default_rng(7)
    p, x = random_problem(rng, 8, 8, M=10, nfcv=10, namp=0, ncp=0, nca=0, regs=())
    V = O.dft_fwd(O.image_1d_to_2d(x, p.xidx, p.yidx, 8, 8), p.u, p.v)
    Vneg = O.dft_fwd(O.image_1d_to_2d(x, p.xidx, p.yidx, 8, 8), -p.u, -p.v)
    assert np.allclose(np.conj(V), Vneg)


# ----------------------------------------------------------------------------
# regularizers
# ----------------------------------------------------------------------------
def _loop_tv_tsv(I2d):
    """image.f90:477-754 restated with 1-based pixel loops (tv_e, tv_grade,
    tsv_e, tsv_grade)."""
    Ny, Nx = I2d.shape

    def I(i, j):
        return I2d[j - 1, i - 1]

    tv_e = np.zeros((Ny, Nx))
    tv_g = np.zeros((Ny, Nx))
    tsv_e = np.zeros((Ny, Nx))
    tsv_g = np.zeros((Ny, Nx))
    for j1 in range(1, Ny + 1):
        for i1 in range(1, Nx + 1):
            i0, j0, i2, j2 = i1 - 1, j1 - 1, i1 + 1, j1 + 1
            dIx = 0.0 if i2 > Nx else I(i2, j1) - I(i1, j1)
            dIy = 0.0 if j2 > Ny else I(i1, j2) - I(i1, j1)
            tv_e[j1 - 1, i1 - 1] = math.sqrt(dIx * dIx + dIy * dIy)
            tsv_e[j1 - 1, i1 - 1] = dIx * dIx + dIy * dIy
            g = 0.0
            e = math.sqrt(dIx * dIx + dIy * dIy)
            if e > O.ZEROEPS:
                g -= (dIx + dIy) / e
            if i0 > 0:
                dIx = I(i1, j1) - I(i0, j1)
                dIy = 0.0 if j2 > Ny else I(i0, j2) - I(i0, j1)
                e = math.sqrt(dIx * dIx + dIy * dIy)
                g = g + dIx / e if e > O.ZEROEPS else 0.0
            if j0 > 0:
                dIy = I(i1, j1) - I(i1, j0)
                dIx = 0.0 if i2 > Nx else I(i2, j0) - I(i1, j0)
                e = math.sqrt(dIx * dIx + dIy * dIy)
                g = g + dIy / e if e > O.ZEROEPS else 0.0
            tv_g[j1 - 1, i1 - 1] = g
            g = 0.0
            if i2 <= Nx:
                g -= 2 * (I(i2, j1) - I(i1, j1))
            if j2 <= Ny:
                g -= 2 * (I(i1, j2) - I(i1, j1))
            if i0 > 0:
                g += 2 * (I(i1, j1) - I(i0, j1))
            if j0 > 0:
                g += 2 * (I(i1, j1) - I(i1, j0))
            tsv_g[j1 - 1, i1 - 1] = g
    return tv_e, tv_g, tsv_e, tsv_g


def test_stencils_match_literal_loops():
    rng = np.random.default_rng(8)
    I2d = rng.standard_normal((5, 7))
    I2d[1:3, 2:4] = 0.4      # flat patch exercises the reset
    tv_e, tv_g, tsv_e, tsv_g = _loop_tv_tsv(I2d)
    assert np.allclose(O.tv_e_map(I2d), tv_e, rtol=0, atol=1e-14)
    assert np.allclose(O.tv_grade_map(I2d), tv_g, rtol=0, atol=1e-13)
    assert np.allclose(O.tsv_e_map(I2d), tsv_e, rtol=0, atol=1e-14)
    assert np.allclose(O.tsv_grade_map(I2d), tsv_g, rtol=0, atol=1e-13)


def test_tv_reset_quirk():
    """Inside a flat 3x3 patch every neighbour term vanishes and the
    reference's `else tv_grade = 0d0` (image.f90:577-578, 603-604) zeroes the
    gradient, although the smooth-TV limit would not."""
    I2d = np.zeros((6, 6))
    I2d[:, :] = np.arange(36).reshape(6, 6) * 0.1
    I2d[1:4, 1:4] = 1.0
    g = O.tv_grade_map(I2d)
    assert g[2, 2] == 0.0 and g[2, 3] == 0.0 and g[3, 2] == 0.0


@pytest.mark.parametrize("which", ["l1", "tsv", "kl", "gs"])
def test_regularizer_gradient_finite_differences(which):
    rng = np.random.default_rng(30)
    p, x = random_problem(rng, 7, 6, M=10, nfcv=5, namp=0, ncp=0, nca=0, regs=(which,))
    if which == "tsv":
        p.tsv_wgt = np.full_like(p.tsv_wgt, 0.7)   # see test_stencil_weight_quirk
    _fd_grad_check(p, x, which="reg", rtol=1e-6)


def test_stencil_weight_quirk():
    """TV / TSV cost and gradient are both weighted by the weight of the
    pixel the loop is at (image.f90:175-182), so with non-uniform weights the
    reference's gradient is not the derivative of its cost.  The oracle keeps
    that: gradient == w_i * tsv_grade_i, which differs from finite
    differences."""
    rng = np.random.default_rng(33)
    p, x = random_problem(rng, 6, 5, M=10, nfcv=5, namp=0, ncp=0, nca=0, regs=("tsv",))
    g = O.calc_cost_reg(x, p)[1]
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, 6, 5)
    lit = p.tsv_l * p.tsv_wgt * _loop_tv_tsv(I2d)[3][p.yidx - 1, p.xidx - 1]
    assert np.allclose(g, lit, rtol=1e-13, atol=0)
    d = rng.standard_normal(x.size)
    h = 1e-6
    fd = (O.calc_cost_reg(x + h * d, p)[0] - O.calc_cost_reg(x - h * d, p)[0]) / (2 * h)
    assert abs(fd - g @ d) > 1e-3 * abs(fd)


def test_tv_gradient_finite_differences_away_from_flat():
    rng = np.random.default_rng(31)
    p, x = random_problem(rng, 6, 6, M=10, nfcv=5, namp=0, ncp=0, nca=0, regs=("tv",))
    p.tv_wgt = np.ones_like(p.tv_wgt)      # the reference's gradient uses the current pixel's weight
    _fd_grad_check(p, x, which="reg", rtol=1e-5)


@pytest.mark.parametrize("which", ["tfd", "cen", "sm"])
def test_moment_regularizer_gradients(which):
    rng = np.random.default_rng(32)
    p, x = random_problem(rng, 8, 7, M=10, nfcv=5, namp=0, ncp=0, nca=0, regs=(), extra_regs=(which,))
    _fd_grad_check(p, x, which="reg", rtol=1e-5)


def test_weight_initialisers_normalisation():
    rng = np.random.default_rng(9)
    prior = rng.uniform(0.1, 1, 30)
    w = O.init_l1reg(prior)
    assert abs(np.sum(w * O.smabs(prior)) - 1.0) < 1e-12          # image.f90:221
    lam, P = O.init_klreg(2.0, prior)
    assert abs(lam * np.sum(O.kl_e(P, P)) - 2.0) < 1e-12          # image.f90:290
    assert abs(np.sum(O.kl_e(P, P)) - np.sum(P) / math.e) < 1e-8   # kl(P,P) = |P|_1/e up to smabs
    xi = np.tile(np.arange(1, 7), 5)
    yi = np.repeat(np.arange(1, 6), 6)
    wt = O.init_tsvreg(xi, yi, 6, 5, False, prior)
    I2d = O.image_1d_to_2d(prior, xi, yi, 6, 5)
    assert abs(np.sum(O.tsv_e_map(I2d)[yi - 1, xi - 1] * wt) - 1.0) < 1e-12
