"""The autotuner on the GPU: every candidate it times reproduces the
default's cost, and the setting it returns is one the library accepts."""
import numpy as np
import pytest

from tests.helpers import problem_kwargs, random_problem

pytestmark = pytest.mark.gpu


def test_autotune_small_problem(gpu_lib):
    from smili_amd.autotune import autotune
    from smili_amd.cost import CostProblem
    rng = np.random.default_rng(61)
    p, x = random_problem(rng, 48, 48, M=800)
    space = (("chunk", (4096, 512)), ("bin", (16, 8)), ("col_group", (0, 1)))
    best, rows = autotune(lambda t: CostProblem(**problem_kwargs(p), tune=t), x, space=space, reps=10)
    assert len(rows) >= 4 and all(r["seconds"] is not None for r in rows)
    with CostProblem(**problem_kwargs(p), tune=best) as g, CostProblem(**problem_kwargs(p)) as d:
        c1, g1, _ = g.cost_grad(x)
        c0, g0, _ = d.cost_grad(x)
    assert abs(c1 - c0) <= 1e-9 * abs(c0)
    assert np.linalg.norm(g1 - g0) <= 1e-9 * np.linalg.norm(g0)
