"""Pins of the oracle's dynamic-imaging cost (fftim3d.calc_cost,
src/fortlib/fftim3d.f90:472-640) that do not restate its formulas:
reductions to the 2-D cost, finite differences, closed forms on constant
and linear-in-time movies, and the light-curve weight rescaling of
calc_cost_reg3d (src/fortlib/image3d.f90:184-236)."""
import dataclasses

import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import random_movie, random_problem


def _split(M, Nz, rng):
    cuts = np.sort(rng.choice(np.arange(1, M), Nz - 1, replace=False))
    return np.diff(np.concatenate([[0], cuts, [M]]))


def test_single_frame_movie_is_the_2d_cost():
    rng = np.random.default_rng(1)
    p, x = random_problem(rng, 12, 10, M=120, extra_regs=("tfd", "sm"))
    mv = O.Movie(frame=p, Nuvs=np.array([p.u.size]))
    c3, g3, t3 = O.calc_cost3d(x, mv)
    c2, g2, t2 = O.calc_cost(x, p)
    assert abs(c3 - c2) <= 1e-13 * abs(c2)
    np.testing.assert_allclose(g3, g2, rtol=1e-12, atol=1e-14 * np.abs(g2).max())
    for k in O.TERM_NAMES:
        assert t3[k] == pytest.approx(t2[k], rel=1e-13, abs=1e-300)


def test_identical_frames_split_uv_reduce_to_2d():
    """Nz identical frames observed on consecutive parts of one uv list:
    the chi-square is the 2-D one, the per-frame chi-square gradients sum to
    the 2-D gradient, the static regularizers average to the 2-D ones
    (gradient / Nz), and Rt / Ri vanish."""
    rng = np.random.default_rng(2)
    Nz = 4
    p, x = random_problem(rng, 14, 11, M=200, window=True)
    nuvs = _split(p.u.size, Nz, rng)
    mv = O.Movie(frame=p, Nuvs=nuvs, rt_l=1.3, rt_wgt=rng.uniform(0.5, 1.5, p.npix),
                 ri_l=0.7, ri_wgt=rng.uniform(0.5, 1.5, p.npix))
    X = np.tile(x, Nz)
    chi3, gchi3, tc3, _ = O.calc_chisq3d(X, mv)
    chi2, gchi2, tc2, *_ = O.calc_chisq(x, p)
    assert abs(chi3 - chi2) <= 1e-12 * abs(chi2)
    np.testing.assert_allclose(gchi3.reshape(Nz, -1).sum(axis=0), gchi2, rtol=1e-9,
                               atol=1e-12 * np.abs(gchi2).max())
    reg3, greg3, tr3 = O.calc_cost_reg3d(X, mv)
    reg2, greg2, _ = O.calc_cost_reg(x, p)
    assert reg3 == pytest.approx(reg2, rel=1e-12)
    np.testing.assert_allclose(greg3.reshape(Nz, -1), np.tile(greg2 / Nz, (Nz, 1)), rtol=1e-12, atol=1e-15)
    assert tr3["rt"] == 0.0 and tr3["ri"] == 0.0


def _dyn_only(mv):
    f = dataclasses.replace(mv.frame, l1_l=-1.0, tv_l=-1.0, tsv_l=-1.0, kl_l=-1.0, gs_l=-1.0)
    return dataclasses.replace(mv, frame=f)


@pytest.mark.parametrize("which", ["rt", "ri"])
def test_temporal_gradients_match_finite_differences(which):
    rng = np.random.default_rng(3)
    Nz = 5
    mv, x = random_movie(rng, Nz, 6, 5, [10] * Nz, rt=which == "rt", ri=which == "ri")
    mv = _dyn_only(mv)
    npix = mv.frame.npix

    def f(xx):
        rt, ri, _ = O.calc_cost_dynamical(xx, mv)
        return rt + ri

    _, _, g = O.calc_cost_dynamical(x, mv)
    h = 1e-6
    fd = np.array([(f(x + h * e) - f(x - h * e)) / (2 * h) for e in np.eye(x.size)])
    if which == "ri":
        # <I> is a function of every frame, but sum_z (I_z - <I>) = 0 makes
        # the reference's constant-<I> gradient exact
        np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)
    else:
        np.testing.assert_allclose(g[:-npix], fd[:-npix], rtol=1e-6, atol=1e-8)
        # the last frame receives no Rt gradient (image3d.f90:374) although
        # Rt depends on it
        assert np.all(g[-npix:] == 0.0)
        assert np.abs(fd[-npix:]).max() > 1e-3


def test_temporal_closed_forms():
    rng = np.random.default_rng(4)
    Nz, npix = 6, 20
    p, _ = random_problem(rng, 5, 4, M=10, nfcv=5, namp=0, ncp=0, nca=0, regs=())
    w_t = rng.uniform(0.5, 1.5, npix)
    w_i = rng.uniform(0.5, 1.5, npix)
    mv = O.Movie(frame=p, Nuvs=np.array([0] * (Nz - 1) + [10]), rt_l=0.9, rt_wgt=w_t, ri_l=1.7, ri_wgt=w_i)
    a = rng.uniform(1, 2, npix)
    b = rng.uniform(-0.2, 0.2, npix)
    z = np.arange(Nz)[:, None]
    lin = (a + b * z).reshape(-1)
    rt, ri, _ = O.calc_cost_dynamical(lin, mv)
    assert rt == pytest.approx(0.9 * np.sum(w_t * b ** 2) * (Nz - 1), rel=1e-12)
    zc = np.arange(Nz) - (Nz - 1) / 2.0
    assert ri == pytest.approx(1.7 * np.sum(w_i * b ** 2) * np.sum(zc ** 2), rel=1e-10)
    const = np.tile(a, Nz)
    rt, ri, g = O.calc_cost_dynamical(const, mv)
    assert rt == 0.0 and abs(ri) < 1e-25 and np.abs(g).max() < 1e-12
    # the light-curve normalisation divides each frame by its target:
    # frames proportional to their targets form a constant normalised movie
    lc = rng.uniform(0.5, 2.0, Nz)
    mv.lc_tgtfd = lc
    rt, ri, g = O.calc_cost_dynamical((lc[:, None] * a[None, :]).reshape(-1), mv)
    assert rt < 1e-25 and ri < 1e-25 and np.abs(g).max() < 1e-12


@pytest.mark.parametrize("nidx", [1, 2, 3, -1])
def test_light_curve_mode(nidx):
    """lc_nidx 1 / 3 rescale the l1 weights by tfd_tgtfd / lc_tgtfd[z]
    (image3d.f90:187-191): with l1 alone, frame z's l1 cost is that of the
    plain weights times the factor; the light-curve term is
    sum_z lc_l[z] (sum x_z - lc_tgtfd[z])^2 / Nz and the tfd term is off."""
    rng = np.random.default_rng(10 + nidx)
    Nz = 4
    mv, x = random_movie(rng, Nz, 8, 6, [15] * Nz, regs=("l1",), rt=False, ri=False, lc=True, lc_nidx=nidx)
    mv.frame.tfd_l = 5.0          # ignored in light-curve mode
    npix = mv.frame.npix
    X = x.reshape(Nz, npix)
    _, _, t = O.calc_cost_reg3d(x, mv)
    lc, F = mv.lc, mv.frame.tfd_tgtfd
    scale = F / lc if nidx in (1, 3) else np.ones(Nz)
    sa = np.sqrt(X ** 2 + 1e-10)               # smabs, image.f90:1146-1158
    l1 = sum(mv.frame.l1_l * np.sum(mv.frame.l1_wgt * scale[z] * sa[z]) for z in range(Nz)) / Nz
    assert t["l1"] == pytest.approx(l1, rel=1e-12)
    lcc = sum(mv.lc_l[z] * (X[z].sum() - lc[z]) ** 2 for z in range(Nz)) / Nz
    assert t["lc"] == pytest.approx(lcc, rel=1e-12)
    assert t["tfd"] == 0.0


def test_centroid_has_no_effect_on_a_movie():
    rng = np.random.default_rng(6)
    mv, x = random_movie(rng, 3, 8, 8, [20, 20, 20])
    c0, g0, _ = O.calc_cost3d(x, mv)
    mv.frame.cen_l, mv.frame.cen_alpha = 3.0, 2.0
    c1, g1, t1 = O.calc_cost3d(x, mv)
    assert c1 == c0 and np.array_equal(g1, g0) and t1["cen"] == 0.0
