"""Full-size (256x256, 1e5 amplitudes + 605,000 closures) GPU evaluation:
checked against the FP64 ES-kernel NUFFT oracle (itself pinned to the exact
DFT at small sizes) and through size-independent properties."""
import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    from smili_amd import workloads
    from smili_amd.cost import CostProblem
    prep, _ = workloads.single_large(seed=0)
    g = CostProblem(**prep.kwargs, tune=dict(precision=32))
    yield prep, g
    g.close()


def test_full_size_matches_fp64_oracle(gpu_lib, big):
    prep, g = big
    x = prep.x0
    c, gr, t = g.cost_grad(x)
    p = O.Problem(**prep.kwargs)
    c_ref, g_ref, t_ref = O.calc_cost(x, p, fwd=O.es_nufft_type2, adj=O.es_nufft_type1)
    print("full size: cost", c, c_ref, abs(c - c_ref) / abs(c_ref), "grad", rel_l2(gr, g_ref))
    for k in ("chisq_amp", "chisq_cp", "chisq_ca", "l1", "tv", "tsv", "kl"):
        print(k, t[k], t_ref[k], abs(t[k] - t_ref[k]) / abs(t_ref[k]))
    assert abs(c - c_ref) <= 2e-6 * abs(c_ref)
    assert rel_l2(gr, g_ref) <= 1e-5


def test_full_size_directional_derivative(gpu_lib, big):
    """(f(x + h d) - f(x - h d)) / 2h == grad . d for a smooth direction,
    on the default FP64 path: with h = 2e-3 the central difference's
    truncation error is ~1e-5 relative (it is ~1e-3 at h = 2e-2, measured)
    and FP64 rounding of the cost adds ~1e-11, so the tolerance is 1e-4.
    (FP32 rounding of the cost, ~3e-8 relative, would swamp the difference
    at any h that keeps the truncation small.)"""
    from smili_amd.cost import CostProblem
    prep, _ = big
    rng = np.random.default_rng(0)
    # no flat patches: TV is not differentiable there (and the reference
    # resets its gradient, image.f90:577-578)
    x = prep.x0 + 0.05 + 0.01 * rng.uniform(size=prep.x0.size)
    d = rng.standard_normal(x.size) * 0.01
    with CostProblem(**prep.kwargs) as g:
        c0, gr, _ = g.cost_grad(x)
        h = 2e-3
        cp, _, _ = g.cost_grad(x + h * d)
        cm, _, _ = g.cost_grad(x - h * d)
    fd = (cp - cm) / (2 * h)
    an = float(gr @ d)
    print("directional", fd, an)
    assert abs(fd - an) <= 1e-4 * abs(an)


def test_full_size_repeatable_cost(gpu_lib, big):
    prep, g = big
    c1, g1, _ = g.cost_grad(prep.x0)
    c2, g2, _ = g.cost_grad(prep.x0)
    assert c1 == c2
    assert rel_l2(g2, g1) < 1e-6


def test_full_size_fp64_matches_fp64_oracle(gpu_lib, big):
    """Default (FP64) NUFFT at full size: same arithmetic as the FP64 ES
    oracle, agreement to rounding."""
    from smili_amd.cost import CostProblem
    prep, _ = big
    p = O.Problem(**prep.kwargs)
    with CostProblem(**prep.kwargs) as g64:
        c, gr, t = g64.cost_grad(prep.x0)
    c_ref, g_ref, _ = O.calc_cost(prep.x0, p, fwd=O.es_nufft_type2, adj=O.es_nufft_type1)
    print("full size fp64: cost", abs(c - c_ref) / abs(c_ref), "grad", rel_l2(gr, g_ref))
    assert abs(c - c_ref) <= 1e-12 * abs(c_ref)
    assert rel_l2(gr, g_ref) <= 1e-10
