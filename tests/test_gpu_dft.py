"""GPU parity of the pruned-DFT grid transforms (smili_tune.transform = 2:
the type-2 / type-1 NUFFT's grid passes as dense products on the window of
fine-grid cells the points' taps touch) against the FP64 ES-kernel oracle
(numpy, full-grid FFTs) and against the FFT path (transform = 1).

Both paths evaluate the same ES-kernel NUFFT (finufft2d2 / finufft2d1 with
the +i sign, src/wrapfinufft/wrapfinufft.f90:94-206), so in FP64 they agree
with the FP64 ES oracle to rounding: visibilities and adjoint rel-l2 <=
1e-11, cost rel <= 1e-12, gradient rel-l2 <= 1e-10 (the tolerances of the
FFT path's FP64 tests); FP32 within the FP32 path's tolerances.  Cases
cover small windows (uv points within a quarter of the Nyquist range, the
usual VLBI case), nearly full coverage, points beyond one period (the
window folds), windowed images, odd sizes, movies and batches that mix
both paths.
"""
import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import movie_kwargs, problem_kwargs, random_movie, random_problem, rel_l2

pytestmark = pytest.mark.gpu

ES = dict(fwd=O.es_nufft_type2, adj=O.es_nufft_type1)


def _mk(p, **tune):
    from smili_amd.cost import CostProblem
    return CostProblem(**problem_kwargs(p), tune=tune)


CASES = [((24, 20), 0.25, False), ((17, 13), 0.3, True), ((64, 48), 0.2, False),
         ((40, 30), 0.95, False), ((33, 17), 0.95, True), ((128, 96), 0.15, False)]


@pytest.mark.parametrize("transform,dft_shapes", [(1, 0), (2, 1), (2, 2), (2, 3)])
@pytest.mark.parametrize("shape,ufrac,window", CASES)
def test_fp64_matches_es_oracle(gpu_lib, shape, ufrac, window, transform, dft_shapes):
    """Both transform paths; the DFT path with both work-group shape sets of
    its products (1: few products, 2: many)."""
    rng = np.random.default_rng(500 + shape[0] + 3 * shape[1] + window)
    p, x = random_problem(rng, shape[0], shape[1], M=600, window=window, umax=np.pi * ufrac)
    a = rng.standard_normal(p.u.size)
    b = rng.standard_normal(p.u.size)
    with _mk(p, precision=64, transform=transform, dft_shapes=dft_shapes) as g:
        V = g.model_vis(x)
        c, gr, t = g.cost_grad(x)
        # the adjoint right after an evaluation, then the evaluation again:
        # the spreading window must be clean for each
        adj = g.adjoint_resid(a, b)
        c2, gr2, _ = g.cost_grad(x)
        prof = g.profile(reps=1)
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny)
    Ve = O.es_nufft_type2(I2d, p.u, p.v)
    adj_e = O.nufft_adj_resid(a, b, p.u, p.v, p.Nx, p.Ny, adj=O.es_nufft_type1)[p.yidx - 1, p.xidx - 1]
    c_es, g_es, t_es = O.calc_cost(x, p, **ES)
    print(shape, ufrac, window, transform, "V", rel_l2(V, Ve), "adj", rel_l2(adj, adj_e),
          "cost", abs(c - c_es) / abs(c_es), "grad", rel_l2(gr, g_es))
    # the requested path ran
    ran_dft = prof["dft2_y"] > 0 and prof["dft1_y"] > 0
    ran_fft = prof["t2_rows"] > 0 and prof["t1_rows"] > 0
    assert (ran_dft, ran_fft) == ((True, False) if transform == 2 else (False, True))
    assert rel_l2(V, Ve) <= 1e-11
    assert rel_l2(adj, adj_e) <= 1e-11
    assert abs(c - c_es) <= 1e-12 * abs(c_es)
    assert rel_l2(gr, g_es) <= 1e-10
    assert c2 == c and rel_l2(gr2, gr) <= 1e-13
    for k, v in t_es.items():
        assert t[k] == pytest.approx(v, rel=1e-10, abs=1e-12 * abs(c_es)), k


@pytest.mark.parametrize("dft_shapes", [1, 2, 3])
@pytest.mark.parametrize("shape,ufrac,window", CASES[:4])
def test_fp32_matches_es_oracle(gpu_lib, shape, ufrac, window, dft_shapes):
    """FP32 products (FP32 DFT matrices, FP32 accumulation over <= 2 nf terms)
    within the FP32 path's tolerances: visibilities rel-l2 <= 1e-6, cost rel
    <= 1e-6, gradient rel-l2 <= 5e-6 + 0.5 e_es."""
    rng = np.random.default_rng(600 + shape[0])
    p, x = random_problem(rng, shape[0], shape[1], M=500, window=window, umax=np.pi * ufrac)
    with _mk(p, precision=32, transform=2, dft_shapes=dft_shapes) as g:
        V = g.model_vis(x)
        c, gr, _ = g.cost_grad(x)
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny)
    Ve = O.es_nufft_type2(I2d, p.u, p.v)
    c_es, g_es, _ = O.calc_cost(x, p, **ES)
    c_d, g_d, _ = O.calc_cost(x, p)
    e_es = rel_l2(g_es, g_d)
    print("fp32", shape, "V", rel_l2(V, Ve), "cost", abs(c - c_es) / abs(c_es), "grad", rel_l2(gr, g_es), e_es)
    assert rel_l2(V, Ve) <= 1e-6
    assert abs(c - c_es) <= 1e-6 * abs(c_es)
    assert rel_l2(gr, g_es) <= 5e-6 + 0.5 * e_es


def test_dft_and_fft_paths_agree(gpu_lib):
    rng = np.random.default_rng(77)
    p, x = random_problem(rng, 48, 40, M=900, umax=np.pi * 0.4, window=True)
    with _mk(p, precision=64, transform=1) as f, _mk(p, precision=64, transform=2) as d:
        cf, gf, tf = f.cost_grad(x)
        cd, gd, td = d.cost_grad(x)
        Vf = f.model_vis(x)
        Vd = d.model_vis(x)
    assert rel_l2(Vd, Vf) <= 1e-12
    assert abs(cd - cf) <= 1e-13 * abs(cf)
    assert rel_l2(gd, gf) <= 1e-11


def test_points_beyond_one_period(gpu_lib):
    """Points at +-pi and beyond the principal period: the window folds the
    coordinates into one period (cells c and c + nf of the window are one
    grid cell, the DFT phases are periodic)."""
    rng = np.random.default_rng(3)
    p, x = random_problem(rng, 16, 16, M=80)
    p.u[:10] = np.pi
    p.u[10:20] = -np.pi
    p.v[20:30] = np.pi
    p.u[30:40] = 2.5 * np.pi
    p.v[40:50] = 3.1 * np.pi
    p.u[50:60] = -2.9 * np.pi
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, 16, 16)
    with _mk(p, precision=64, transform=2) as g:
        V = g.model_vis(x)
        c, gr, _ = g.cost_grad(x)
    assert rel_l2(V, O.es_nufft_type2(I2d, p.u, p.v)) <= 1e-11
    assert rel_l2(V, O.dft_fwd(I2d, p.u, p.v)) <= 4e-6
    c_es, g_es, _ = O.calc_cost(x, p, **ES)
    assert abs(c - c_es) <= 1e-12 * abs(c_es)
    assert rel_l2(gr, g_es) <= 1e-10


def test_small_window_is_the_default(gpu_lib):
    """transform = 0 takes the products for a small window and the FFTs
    for full coverage."""
    rng = np.random.default_rng(8)
    small, xs = random_problem(rng, 64, 64, M=400, umax=np.pi * 0.2)
    full, xf = random_problem(rng, 64, 64, M=400, umax=np.pi * 0.95)
    with _mk(small) as g:
        ps = g.profile(reps=1)
    with _mk(full) as g:
        pf = g.profile(reps=1)
    assert ps["dft2_y"] > 0 and ps["t2_rows"] == 0
    assert pf["t2_rows"] > 0 and pf["dft2_y"] == 0


def test_batch_mixing_both_paths(gpu_lib):
    """A batch of small-window and full-coverage problems runs both paths in
    one plan; each problem matches its own handle and the oracle."""
    from smili_amd.cost import BatchProblem
    rng = np.random.default_rng(13)
    cases = [random_problem(rng, 24, 20, M=200, umax=np.pi * 0.25),
             random_problem(rng, 64, 64, M=300, umax=np.pi * 0.95, window=True),
             random_problem(rng, 40, 30, M=250, umax=np.pi * 0.15, nfcv=0),
             random_problem(rng, 17, 13, M=150, umax=np.pi * 0.9)]
    X = np.concatenate([x for _, x in cases])
    with BatchProblem([problem_kwargs(p) for p, _ in cases]) as b:
        prof = b.profile(reps=1)
        costs, grad, _ = b.cost_grad(X)
        grads = b.split(grad)
    assert prof["dft2_y"] > 0 and prof["t2_rows"] > 0
    for i, (p, x) in enumerate(cases):
        with _mk(p) as s:
            c1, g1, _ = s.cost_grad(x)
        c_es, g_es, _ = O.calc_cost(x, p, **ES)
        assert abs(costs[i] - c1) <= 1e-13 * abs(c1)
        assert rel_l2(grads[i], g1) <= 1e-12
        assert abs(costs[i] - c_es) <= 1e-11 * abs(c_es)
        assert rel_l2(grads[i], g_es) <= 1e-10


@pytest.mark.parametrize("dft_shapes", [1, 2, 3])
def test_movie_on_small_windows(gpu_lib, dft_shapes):
    from smili_amd.cost import MovieProblem
    rng = np.random.default_rng(14)
    mv, x = random_movie(rng, 5, 32, 24, [150, 0, 120, 200, 90], umax=np.pi * 0.3, window=True)
    with MovieProblem(**movie_kwargs(mv), tune={"transform": 2, "dft_shapes": dft_shapes}) as g:
        cost, grad, terms = g.cost_grad(x)
        prof = g.profile(reps=1)
    c_es, g_es, t_es = O.calc_cost3d(x, mv, **ES)
    assert prof["dft2_y"] > 0 and prof["t2_rows"] == 0
    assert abs(cost - c_es) <= 1e-11 * abs(c_es)
    assert rel_l2(grad, g_es) <= 1e-10


def test_grids_beyond_the_lds_fft_take_the_products(gpu_lib):
    """A 768 x 640 image has a 1536 x 1280 fine grid, whose rows and columns
    do not fit the LDS FFT's 64 KB in FP64: the default takes the products
    even at nearly full coverage."""
    rng = np.random.default_rng(15)
    p, x = random_problem(rng, 768, 640, M=300, umax=np.pi * 0.9, nfcv=150, namp=100, ncp=60, nca=50)
    with _mk(p) as g:
        prof = g.profile(reps=1)
        c, gr, _ = g.cost_grad(x)
    assert prof["dft2_y"] > 0 and prof["t2_rows"] == 0
    c_es, g_es, _ = O.calc_cost(x, p, **ES)
    assert abs(c - c_es) <= 1e-12 * abs(c_es)
    assert rel_l2(gr, g_es) <= 1e-10
