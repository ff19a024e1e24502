"""GPU parity of the HIP cost against the FP64 oracle, through the C ABI.

Tolerances (FP32 spreading / interpolation / FFT with an ES kernel of width
7 at eps 1e-6, FP64 chi-square and regularizer accumulation):
  * type-2 model visibilities vs the exact DFT: rel-l2 <= 4e-6
  * type-1 adjoint vs the exact DFT: rel-l2 <= 4e-6
  * cost vs the exact-DFT cost: rel <= 2e-6; each chi-square term rel <= 2e-5
  * gradient vs the exact-DFT gradient: rel-l2 <= 1e-5
  * regularizer terms (FP64 end to end): rel <= 1e-12
"""
import math

import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import problem_kwargs, random_problem, rel_l2

pytestmark = pytest.mark.gpu

TOL_V = 4e-6
TOL_COST = 2e-6
TOL_TERM = 2e-5
TOL_GRAD = 1e-5
TOL_REG = 1e-12


def _mk(p, **kw):
    from smili_amd.cost import CostProblem
    return CostProblem(**problem_kwargs(p), **kw)


SHAPES = [(16, 16), (24, 20), (17, 13), (32, 48), (64, 64), (40, 30)]


@pytest.mark.parametrize("shape", SHAPES)
def test_type2_matches_dft(gpu_lib, shape):
    rng = np.random.default_rng(100 + shape[0])
    p, x = random_problem(rng, shape[0], shape[1], M=400)
    with _mk(p) as g:
        V = g.model_vis(x)
    Vd = O.dft_fwd(O.image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny), p.u, p.v)
    err = rel_l2(V, Vd)
    print("type2", shape, err)
    assert err <= TOL_V


@pytest.mark.parametrize("shape", SHAPES)
def test_type1_matches_dft(gpu_lib, shape):
    rng = np.random.default_rng(200 + shape[0])
    p, x = random_problem(rng, shape[0], shape[1], M=400)
    a = rng.standard_normal(p.u.size)
    b = rng.standard_normal(p.u.size)
    with _mk(p) as g:
        out = g.adjoint_resid(a, b)
    ref = O.nufft_adj_resid(a, b, p.u, p.v, p.Nx, p.Ny)[p.yidx - 1, p.xidx - 1]
    err = rel_l2(out, ref)
    print("type1", shape, err)
    assert err <= TOL_V


def _check_cost(p, x, g, tol_cost=TOL_COST, tol_grad=TOL_GRAD):
    c_ref, g_ref, t_ref = O.calc_cost(x, p)
    c, gr, t = g.cost_grad(x)
    errs = {k: abs(t[k] - t_ref[k]) / max(abs(t_ref[k]), 1e-300) for k in O.TERM_NAMES if t_ref[k] != 0}
    print("cost", abs(c - c_ref) / abs(c_ref), "grad", rel_l2(gr, g_ref), errs)
    assert abs(c - c_ref) <= tol_cost * abs(c_ref)
    for k in ("chisq_fcv", "chisq_amp", "chisq_cp", "chisq_ca"):
        if t_ref[k] != 0:
            assert abs(t[k] - t_ref[k]) <= TOL_TERM * abs(t_ref[k]), k
    for k in ("l1", "tv", "tsv", "kl", "gs", "tfd", "cen", "sm"):
        if t_ref[k] != 0:
            assert abs(t[k] - t_ref[k]) <= TOL_REG * abs(t_ref[k]) + 1e-300, k
        else:
            assert t[k] == 0.0, k
    assert rel_l2(gr, g_ref) <= tol_grad


@pytest.mark.parametrize("shape,window", [((24, 20), False), ((32, 32), True), ((17, 13), False),
                                          ((64, 48), True)])
def test_cost_grad_all_terms(gpu_lib, shape, window):
    rng = np.random.default_rng(7 + shape[0] + window)
    p, x = random_problem(rng, shape[0], shape[1], M=500, window=window)
    with _mk(p) as g:
        _check_cost(p, x, g)


@pytest.mark.parametrize("only", ["fcv", "amp", "cp", "ca"])
def test_single_data_type(gpu_lib, only):
    rng = np.random.default_rng({"fcv": 1, "amp": 2, "cp": 3, "ca": 4}[only])
    n = dict(nfcv=0, namp=0, ncp=0, nca=0)
    n["n" + only] = 300
    p, x = random_problem(rng, 20, 20, M=250, regs=(), **n)
    with _mk(p) as g:
        _check_cost(p, x, g)


def test_moment_regularizers(gpu_lib):
    rng = np.random.default_rng(11)
    p, x = random_problem(rng, 24, 24, M=200, extra_regs=("tfd", "cen", "sm"))
    with _mk(p) as g:
        _check_cost(p, x, g)


def test_tv_reset_quirk_on_flat_patch(gpu_lib):
    """A flat patch makes neighbour TV terms vanish: the reference resets
    the accumulated gradient to 0 there (image.f90:577-578, 603-604)."""
    rng = np.random.default_rng(5)
    p, x = random_problem(rng, 16, 16, M=100, regs=("tv",))
    x2 = O.image_1d_to_2d(x, p.xidx, p.yidx, 16, 16)
    x2[4:9, 4:9] = 1.0
    x = x2[p.yidx - 1, p.xidx - 1]
    with _mk(p) as g:
        _check_cost(p, x, g)


def test_graph_and_eager_agree_and_repeat(gpu_lib):
    rng = np.random.default_rng(9)
    p, x = random_problem(rng, 32, 32, M=600)
    with _mk(p) as g1, _mk(p, tune=dict(use_graph=-1)) as g2:
        c1, gr1, _ = g1.cost_grad(x)
        c1b, gr1b, _ = g1.cost_grad(x)
        c2, gr2, _ = g2.cost_grad(x)
        assert g1.info()["graph"] == 1 and g2.info()["graph"] == 0
    # the cost is a fixed-order reduction: bitwise reproducible
    assert c1 == c1b == c2
    # the gradient goes through float atomics: reproducible to rounding
    assert rel_l2(gr1b, gr1) < 1e-6 and rel_l2(gr2, gr1) < 1e-6


@pytest.mark.parametrize("tune", [dict(bin=8, sub_points=64), dict(bin=32, sub_points=1024, col_group=8),
                                  dict(chunk=64)])
def test_tunables_do_not_change_results(gpu_lib, tune):
    rng = np.random.default_rng(21)
    p, x = random_problem(rng, 40, 40, M=700)
    with _mk(p) as g, _mk(p, tune=tune) as h:
        c1, gr1, _ = g.cost_grad(x)
        c2, gr2, _ = h.cost_grad(x)
    assert abs(c1 - c2) <= 1e-6 * abs(c1)
    assert rel_l2(gr2, gr1) < 1e-6


def test_points_on_the_period_boundary(gpu_lib):
    """uv points at exactly +-pi and beyond the principal period wrap
    (the transform is 2 pi periodic in u' for integer modes)."""
    rng = np.random.default_rng(3)
    p, x = random_problem(rng, 16, 16, M=50)
    p.u[:10] = np.pi
    p.u[10:20] = -np.pi
    p.v[20:30] = np.pi
    p.u[30:40] = 2.5 * np.pi
    with _mk(p) as g:
        V = g.model_vis(x)
    Vd = O.dft_fwd(O.image_1d_to_2d(x, p.xidx, p.yidx, 16, 16), p.u, p.v)
    assert rel_l2(V, Vd) <= TOL_V


def test_nonfinite_is_reported(gpu_lib):
    """A zero image gives |V| = 0 and the amplitude term divides by it
    (fftlib.f90:488): the reference returns NaN silently, the ABI flags it."""
    from smili_amd._lib import SmiliError
    rng = np.random.default_rng(4)
    p, x = random_problem(rng, 16, 16, M=50, nfcv=0, ncp=0, nca=0, regs=())
    with _mk(p) as g:
        with pytest.raises(SmiliError) as e:
            g.cost_grad(np.zeros_like(x))
        assert e.value.code == -4
        c, gr, t = g.cost_grad(np.zeros_like(x), allow_nonfinite=True)
        assert not math.isfinite(c)
