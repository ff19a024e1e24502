"""The batched survey (512 problems of 64^2 .. 256^2 with 2e3 .. 5e4
visibilities each, the benchmark's survey workload) and the 64-frame movie
(128^2 frames, 3e5 visibilities, the benchmark's movie workload) at their
full sizes on the GPU:

* every one of the 512 problems of the batch against its own single handle
  (cost rel <= 1e-13, gradient rel-l2 <= 1e-12: the same kernels on the same
  data, up to the spreading atomics' rounding), one problem of each image
  size against the FP64 ES-kernel oracle (cost 1e-11, gradient 1e-10), the
  captured graph against eager launches, and a directional derivative of
  the summed cost over the whole batch;
* the movie against the oracle's calc_cost3d with the ES transforms (cost
  1e-11, gradient 1e-10, every term) and a directional derivative along a
  direction that keeps the last frame fixed (the reference gives the last
  frame no Rt gradient, image3d.f90:374, and so does this build).
The directional derivatives use h = 1e-3 relative steps, whose central
differences carry a truncation error ~1e-6 (the costs are smooth in the
positive orthant: L1 through smabs, TSV, KL, chi-square); tolerance 1e-4."""
import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

ES = dict(fwd=O.es_nufft_type2, adj=O.es_nufft_type1)


@pytest.fixture(scope="module")
def survey():
    from smili_amd import workloads
    from smili_amd.cost import BatchProblem
    preps, meta = workloads.survey(seed=3)
    kws = [p.kwargs for p in preps]
    X = np.concatenate([p.x0 for p in preps])
    b = BatchProblem(kws)
    yield preps, meta, kws, X, b
    b.close()


def test_survey_every_problem_matches_its_single_handle(gpu_lib, survey):
    from smili_amd.cost import CostProblem
    preps, meta, kws, X, b = survey
    assert b.nprob == 512
    costs, grad, _ = b.cost_grad(X)
    grads = b.split(grad)
    worst_c = worst_g = 0.0
    for i, p in enumerate(preps):
        with CostProblem(**kws[i], tune={"use_graph": -1}) as s:
            c1, g1, _ = s.cost_grad(p.x0)
        worst_c = max(worst_c, abs(costs[i] - c1) / abs(c1))
        worst_g = max(worst_g, rel_l2(grads[i], g1))
    print("512 problems vs single handles: cost", worst_c, "grad", worst_g)
    assert worst_c <= 1e-13
    assert worst_g <= 1e-12


def test_survey_one_problem_per_size_matches_es_oracle(gpu_lib, survey):
    preps, meta, kws, X, b = survey
    costs, grad, terms = b.cost_grad(X)
    grads = b.split(grad)
    seen = set()
    for i, m in enumerate(meta):
        if m["n"] in seen:
            continue
        seen.add(m["n"])
        p = O.Problem(**kws[i])
        c_es, g_es, t_es = O.calc_cost(preps[i].x0, p, **ES)
        print(m["n"], m["nvis"], "cost", abs(costs[i] - c_es) / abs(c_es), "grad", rel_l2(grads[i], g_es))
        assert abs(costs[i] - c_es) <= 1e-11 * abs(c_es)
        assert rel_l2(grads[i], g_es) <= 1e-10
        np.testing.assert_allclose(terms[i], O.terms_vector(t_es), rtol=1e-10, atol=1e-12 * abs(c_es))
    assert seen == {64, 96, 128, 192, 256}


def test_survey_graph_equals_eager(gpu_lib, survey):
    from smili_amd.cost import BatchProblem
    preps, meta, kws, X, b = survey
    costs_e, grad_e, _ = b.cost_grad(X)
    # a batch of this size (two branches) launches one by one by default
    with BatchProblem(kws, tune={"use_graph": 1}) as gr:
        costs, grad, _ = gr.cost_grad(X)
        assert b.info()["graph"] == 0 and gr.info()["graph"] == 1
    assert np.array_equal(costs, costs_e)
    assert rel_l2(grad_e, grad) <= 1e-13


def test_survey_directional_derivative(gpu_lib, survey):
    preps, meta, kws, X, b = survey
    rng = np.random.default_rng(1)
    x = X * (1.0 + 0.1 * rng.uniform(size=X.size)) + 1e-3 * X.max()
    d = rng.standard_normal(X.size) * x
    h = 1e-3
    c0, g0, _ = b.cost_grad(x)
    cp = np.sum(b.cost_grad(x + h * d)[0])
    cm = np.sum(b.cost_grad(x - h * d)[0])
    fd = (cp - cm) / (2 * h)
    an = float(g0 @ d)
    print("survey directional", fd, an)
    assert abs(fd - an) <= 1e-4 * abs(an)


@pytest.fixture(scope="module")
def movie():
    from smili_amd import workloads
    from smili_amd.cost import MovieProblem
    prep, meta = workloads.dynamic_movie(seed=5)
    g = MovieProblem(**prep.kwargs)
    yield prep, meta, g
    g.close()


def test_movie_full_size_matches_es_oracle(gpu_lib, movie):
    from tests.test_gpu_movie import _oracle_movie
    prep, meta, g = movie
    assert prep.kwargs["Nz"] == 64 and int(np.sum(prep.kwargs["Nuvs"])) == 300_000
    x = prep.x0 * (1.0 + 0.05 * np.random.default_rng(2).uniform(size=prep.x0.size))
    cost, grad, terms = g.cost_grad(x)
    mv = _oracle_movie(prep.kwargs)
    c_es, g_es, t_es = O.calc_cost3d(x, mv, **ES)
    print("movie: cost", abs(cost - c_es) / abs(c_es), "grad", rel_l2(grad, g_es))
    assert abs(cost - c_es) <= 1e-11 * abs(c_es)
    assert rel_l2(grad, g_es) <= 1e-10
    for k, v in t_es.items():
        assert terms[k] == pytest.approx(v, rel=1e-10, abs=1e-12 * abs(c_es)), k


def test_movie_full_size_directional_derivative(gpu_lib, movie):
    prep, meta, g = movie
    rng = np.random.default_rng(3)
    x = prep.x0 * (1.0 + 0.1 * rng.uniform(size=prep.x0.size)) + 1e-3 * prep.x0.max()
    d = rng.standard_normal(x.size) * x
    # the last frame has no Rt gradient (image3d.f90:374, kept): a direction
    # that leaves it fixed
    d[-prep.kwargs["xidx"].size:] = 0.0
    h = 1e-3
    c0, g0, _ = g.cost_grad(x)
    fd = (g.cost_grad(x + h * d)[0] - g.cost_grad(x - h * d)[0]) / (2 * h)
    an = float(g0 @ d)
    print("movie directional", fd, an)
    assert abs(fd - an) <= 1e-4 * abs(an)
