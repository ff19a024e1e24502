"""GPU parity of the dynamic-imaging cost (fftim3d.calc_cost,
src/fortlib/fftim3d.f90:472-640) and of batched 2-D problems, through the
C ABI, against the FP64 oracle.

FP64 NUFFT (the default): against the FP64 ES-kernel oracle (the same
approximation in numpy) cost rel <= 1e-11, gradient rel-l2 <= 1e-10, each
term rel <= 1e-10; temporal / light-curve terms are FP64 end to end.
FP32 NUFFT: cost rel <= 1e-6 against the same oracle, gradient rel-l2 <=
5e-6 + 0.5 e_es (e_es: the FP64 ES oracle's own error against the exact DFT,
as in the 2-D parity tests).
"""
import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import movie_kwargs, problem_kwargs, random_movie, random_problem, rel_l2

pytestmark = pytest.mark.gpu

ES = dict(fwd=O.es_nufft_type2, adj=O.es_nufft_type1)


def _movie(mv, tune=None):
    from smili_amd.cost import MovieProblem
    return MovieProblem(**movie_kwargs(mv), tune=tune)


def _check_terms(t_gpu, t_ref, rel):
    for k, v in t_ref.items():
        assert t_gpu[k] == pytest.approx(v, rel=rel, abs=1e-12 * max(abs(t_ref["cost"]), 1.0)), k


CASES = {
    "plain": dict(Nz=5, shape=(24, 20), nuvs=[120, 0, 90, 150, 60], kw={}),
    "window_tfd_sm": dict(Nz=4, shape=(17, 13), nuvs=[80, 70, 60, 90],
                          kw=dict(window=True, extra_regs=("tfd", "sm"))),
    "lc_static": dict(Nz=4, shape=(16, 16), nuvs=[60, 60, 0, 60], kw=dict(lc=True, lc_nidx=1)),
    "lc_dynamic": dict(Nz=3, shape=(20, 18), nuvs=[100, 50, 70], kw=dict(lc=True, lc_nidx=2)),
    "lc_both_window": dict(Nz=3, shape=(20, 18), nuvs=[100, 50, 70], kw=dict(lc=True, lc_nidx=3, window=True)),
    "one_frame": dict(Nz=1, shape=(24, 24), nuvs=[200], kw={}),
    "no_temporal": dict(Nz=6, shape=(12, 12), nuvs=[30] * 6, kw=dict(rt=False, ri=False)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_movie_fp64_matches_es_oracle(gpu_lib, case):
    c = CASES[case]
    rng = np.random.default_rng(20 + sorted(CASES).index(case))
    mv, x = random_movie(rng, c["Nz"], *c["shape"], c["nuvs"], **c["kw"])
    with _movie(mv) as g:
        cost, grad, terms = g.cost_grad(x)
    c_es, g_es, t_es = O.calc_cost3d(x, mv, **ES)
    print(case, "cost", abs(cost - c_es) / abs(c_es), "grad", rel_l2(grad, g_es))
    assert abs(cost - c_es) <= 1e-11 * abs(c_es)
    assert rel_l2(grad, g_es) <= 1e-10
    _check_terms(terms, t_es, 1e-10)
    # temporal and light-curve terms are FP64 arithmetic only
    for k in ("rt", "ri", "lc"):
        assert terms[k] == pytest.approx(t_es[k], rel=1e-12, abs=1e-300)


def test_movie_fp32_matches_es_oracle(gpu_lib):
    rng = np.random.default_rng(7)
    mv, x = random_movie(rng, 4, 32, 32, [300, 200, 0, 250])
    with _movie(mv, tune={"precision": 32}) as g:
        cost, grad, _ = g.cost_grad(x)
    c_es, g_es, _ = O.calc_cost3d(x, mv, **ES)
    c_d, g_d, _ = O.calc_cost3d(x, mv)
    e_es = rel_l2(g_es, g_d)
    print("fp32 movie cost", abs(cost - c_es) / abs(c_es), "grad", rel_l2(grad, g_es), "e_es", e_es)
    assert abs(cost - c_es) <= 1e-6 * abs(c_es)
    assert rel_l2(grad, g_es) <= 5e-6 + 0.5 * e_es


def test_movie_centroid_is_ignored(gpu_lib):
    rng = np.random.default_rng(8)
    mv, x = random_movie(rng, 3, 16, 16, [50, 60, 70])
    with _movie(mv) as g:
        c0, g0, _ = g.cost_grad(x)
    mv.frame.cen_l, mv.frame.cen_alpha = 2.0, 1.0
    with _movie(mv) as g:
        c1, g1, t1 = g.cost_grad(x)
    assert c1 == c0 and t1["cen"] == 0.0
    assert rel_l2(g1, g0) <= 1e-14


def test_movie_transforms_per_frame(gpu_lib):
    """model_vis / adjoint_resid of a movie are the per-frame transforms of
    fftlib3d.f90:76-90 and 126-142, frames in uv-list order."""
    rng = np.random.default_rng(9)
    nuvs = [70, 0, 40, 90]
    mv, x = random_movie(rng, 4, 20, 16, nuvs, window=True)
    p = mv.frame
    npix = p.npix
    start = np.concatenate([[0], np.cumsum(nuvs)])
    with _movie(mv) as g:
        V = g.model_vis(x)
        r = rng.standard_normal(p.u.size), rng.standard_normal(p.u.size)
        A = g.adjoint_resid(*r)
    for z in range(4):
        s, e = start[z], start[z + 1]
        A_ref = np.zeros(npix)
        if e > s:
            I2d = O.image_1d_to_2d(x[z * npix:(z + 1) * npix], p.xidx, p.yidx, p.Nx, p.Ny)
            assert rel_l2(V[s:e], O.es_nufft_type2(I2d, p.u[s:e], p.v[s:e])) <= 1e-12
            g2 = O.nufft_adj_resid(r[0][s:e], r[1][s:e], p.u[s:e], p.v[s:e], p.Nx, p.Ny, O.es_nufft_type1)
            A_ref = g2[p.yidx - 1, p.xidx - 1]
            assert rel_l2(A[z * npix:(z + 1) * npix], A_ref) <= 1e-12
        else:
            assert np.all(A[z * npix:(z + 1) * npix] == 0.0)


def test_movie_graph_equals_eager(gpu_lib):
    rng = np.random.default_rng(10)
    mv, x = random_movie(rng, 5, 24, 24, [100] * 5, lc=True)
    with _movie(mv, tune={"use_graph": 1}) as g, _movie(mv) as e:
        for k in range(3):
            xx = x * (1 + 0.05 * k)
            cg, gg, _ = g.cost_grad(xx)
            ce, ge, _ = e.cost_grad(xx)
            assert cg == ce
            assert rel_l2(gg, ge) <= 1e-13


def test_movie_deterministic_mode(gpu_lib):
    """smili_tune.deterministic on a movie: repeated evaluations give the
    same cost and gradient bit for bit, equal to the default path's to
    rounding."""
    rng = np.random.default_rng(12)
    mv, x = random_movie(rng, 5, 24, 24, [400] * 5, lc=True)
    with _movie(mv) as f, _movie(mv, tune={"deterministic": 1}) as d:
        cf, gf, _ = f.cost_grad(x)
        c0, g0, _ = d.cost_grad(x)
        assert abs(c0 - cf) <= 1e-12 * abs(cf)
        assert rel_l2(g0, gf) <= 1e-10
        for _ in range(3):
            c1, g1, _ = d.cost_grad(x)
            assert c1 == c0 and np.array_equal(g1, g0)


# ---------------------------------------------------------------------------
# batched 2-D problems
# ---------------------------------------------------------------------------
def _batch_cases(rng):
    return [
        random_problem(rng, 16, 16, M=150),
        random_problem(rng, 24, 20, M=300, window=True, extra_regs=("tfd", "cen", "sm")),
        random_problem(rng, 17, 13, M=120, nfcv=0),
        random_problem(rng, 32, 48, M=500, ncp=0, nca=0),
        random_problem(rng, 40, 30, M=250, nfcv=0, namp=0),
    ]


def test_batch_matches_single_problems_and_oracle(gpu_lib):
    from smili_amd.cost import BatchProblem, CostProblem
    rng = np.random.default_rng(11)
    cases = _batch_cases(rng)
    with BatchProblem([problem_kwargs(p) for p, _ in cases]) as b:
        assert b.nprob == len(cases)
        X = np.concatenate([x for _, x in cases])
        costs, grad, terms = b.cost_grad(X)
        grads = b.split(grad)
        for i, (p, x) in enumerate(cases):
            with CostProblem(**problem_kwargs(p)) as s:
                c1, g1, t1 = s.cost_grad(x)
            c_es, g_es, t_es = O.calc_cost(x, p, **ES)
            print(i, "cost vs single", abs(costs[i] - c1) / abs(c1), "vs oracle", abs(costs[i] - c_es) / abs(c_es))
            assert abs(costs[i] - c1) <= 1e-13 * abs(c1)
            assert rel_l2(grads[i], g1) <= 1e-12
            assert abs(costs[i] - c_es) <= 1e-11 * abs(c_es)
            assert rel_l2(grads[i], g_es) <= 1e-10
            np.testing.assert_allclose(terms[i], O.terms_vector(t_es), rtol=1e-10,
                                       atol=1e-12 * max(abs(c_es), 1.0))


def test_batch_survey_of_many_small_problems(gpu_lib):
    """64 independent 32x32 problems (a survey) in one handle, graph and
    eager, against the same problems one by one."""
    from smili_amd.cost import BatchProblem, CostProblem
    rng = np.random.default_rng(12)
    cases = [random_problem(rng, 32, 32, M=int(rng.integers(40, 160)), nfcv=60, namp=40, ncp=30, nca=20,
                            regs=("l1", "tsv")) for _ in range(64)]
    X = np.concatenate([x for _, x in cases])
    kws = [problem_kwargs(p) for p, _ in cases]
    with BatchProblem(kws, tune={"use_graph": 1}) as b, BatchProblem(kws, tune={"use_graph": -1}) as be:
        costs, grad, _ = b.cost_grad(X)
        costs_e, grad_e, _ = be.cost_grad(X)
        assert np.array_equal(costs, costs_e)
        assert rel_l2(grad, grad_e) <= 1e-13
        for i in (0, 17, 63):
            p, x = cases[i]
            with CostProblem(**kws[i]) as s:
                c1, g1, _ = s.cost_grad(x)
            assert abs(costs[i] - c1) <= 1e-13 * abs(c1)
            assert rel_l2(b.split(grad)[i], g1) <= 1e-12


# ---------------------------------------------------------------------------
# the dynamic imager end to end
# ---------------------------------------------------------------------------
def _oracle_movie(kw):
    kw = dict(kw)
    m = {k: kw.pop(k) for k in ("Nz", "Nuvs", "rt_l", "rt_wgt", "ri_l", "ri_wgt", "lc_l", "lc_tgtfd", "lc_nidx")}
    lc_l = None if np.sum(m["lc_l"]) <= 0 else m["lc_l"]
    return O.Movie(frame=O.Problem(**kw), Nuvs=m["Nuvs"], rt_l=m["rt_l"], rt_wgt=m["rt_wgt"], ri_l=m["ri_l"],
                   ri_wgt=m["ri_wgt"], lc_l=lc_l, lc_tgtfd=m["lc_tgtfd"], lc_nidx=m["lc_nidx"])


def test_imaging3d_matches_oracle_solve(gpu_lib):
    """imaging3d (GPU cost, host L-BFGS-B) against the same L-BFGS-B driven
    by the FP64 ES-kernel oracle of fftim3d.calc_cost: frames agree to
    NRMSE <= 1e-4 after 30 iterations, final costs to 1e-6 relative."""
    from scipy.optimize import minimize

    from smili_amd import workloads
    from smili_amd.imaging import lbfgs3d
    from smili_amd.imaging.image import Movie
    from smili_amd.imaging.prep3d import prepare3d

    Nt, n = 4, 32
    _, meta = workloads.dynamic_movie(seed=21, n_frames=Nt, n=n, n_vis=1200, n_st=6, per_frame=(200, 400))
    vis = meta["vis"]
    init = synth_init = np.full((Nt, n, n), 0.6 / (n * n))
    regs = dict(l1_lambda=1.0, tsv_lambda=1e3, rt_lambda=1.0, ri_lambda=0.3)
    movie = Movie.from_frames(init, dx=2.0, angunit="uas")
    out, info = lbfgs3d.imaging3d(movie, vistable=vis, niter=30, verbose=False, return_info=True, **regs)
    got = out.frames()

    pm = prepare3d(synth_init, -2.0 * np.pi / 180 / 3600e6, 2.0 * np.pi / 180 / 3600e6, vistable=vis, **regs)
    mv = _oracle_movie(pm.kwargs)

    def fg(x):
        c, g, _ = O.calc_cost3d(x, mv, **ES)
        return c, g

    res = minimize(fg, pm.x0, jac=True, method="L-BFGS-B", bounds=[(0.0, None)] * pm.x0.size,
                   options=dict(maxcor=5, ftol=1e1 * np.finfo(float).eps, gtol=0.0, maxiter=31, maxfun=10 ** 9,
                                maxls=20))
    ref = (res.x / pm.inormfactr).reshape(Nt, n, n)
    err = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))
    print("imaging3d nit", info.niter, "ref nit", res.nit, "NRMSE", err, "cost", info.cost, res.fun)
    assert info.niter == res.nit
    assert err <= 1e-4
    assert abs(info.cost - res.fun) <= 1e-6 * abs(res.fun)
