"""The C + OpenMP FP64 restatement (CPU baseline) agrees with the numpy
FP64 ES-kernel oracle (same algorithm) and, through it, with the exact DFT."""
import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import problem_kwargs, random_problem, rel_l2


@pytest.fixture(scope="module")
def cpu_mod():
    import os
    from oracle import build
    from oracle.cpu import LIB
    if not os.path.exists(LIB):
        build.build_all()
    from oracle import cpu
    return cpu


@pytest.mark.parametrize("shape,window", [((24, 20), False), ((17, 13), True), ((32, 32), False)])
def test_cpu_port_matches_es_oracle(cpu_mod, shape, window):
    rng = np.random.default_rng(400 + shape[0])
    p, x = random_problem(rng, shape[0], shape[1], M=300, window=window)
    cc = cpu_mod.CpuCost(**problem_kwargs(p))
    c, g, t = cc.cost_grad(x)
    ce, ge, te = O.calc_cost(x, p, fwd=O.es_nufft_type2, adj=O.es_nufft_type1)
    assert abs(c - ce) <= 1e-11 * abs(ce)
    assert rel_l2(g, ge) <= 1e-9
    tv = O.terms_vector(te)
    assert np.allclose(t, tv, rtol=1e-10, atol=1e-300)
