"""smili_cost_grad_device_noalloc (the evaluation with its intermediates in
the library's static device scratch) against smili_cost_grad_device on the
same handle, through the C ABI: two problems in a row, the second smaller,
so the second call runs on scratch the first one left dirty.  The cost is a
fixed-order sum (bitwise equal); the gradient goes through float atomics in
the spreading (equal to rounding: rel-l2 <= 1e-13 in FP64, <= 1e-6 with
FP32 grids, the bound the graph-versus-eager test of the FP32 path uses)."""
import numpy as np
import pytest

from tests.helpers import problem_kwargs, random_problem, rel_l2

pytestmark = pytest.mark.gpu


def _both(g, x0):
    import torch
    n = x0.size
    x = torch.tensor(x0, dtype=torch.float64, device="cuda")
    ga = torch.empty(n, dtype=torch.float64, device="cuda")
    gb = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    ta = torch.empty(g.nterms, dtype=torch.float64, device="cuda")
    tb = torch.full((g.nterms,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.cost_grad_device(x.data_ptr(), ga.data_ptr(), ta.data_ptr())
    g.synchronize()
    g.cost_grad_device_noalloc(x.data_ptr(), gb.data_ptr(), tb.data_ptr())
    g.synchronize()
    return ga.cpu().numpy(), ta.cpu().numpy(), gb.cpu().numpy(), tb.cpu().numpy()


def test_noalloc_matches_the_device_entry_point(gpu_lib):
    from smili_amd import workloads
    from smili_amd.cost import CostProblem
    rng = np.random.default_rng(61)
    big, _ = workloads.single_large(seed=0)
    small, xs = random_problem(rng, 24, 20, M=300, window=True, extra_regs=("tfd", "cen", "sm"))
    for kw, x0, tune in ((big.kwargs, big.x0, None), (problem_kwargs(small), xs, {"transform": 1}),
                         (problem_kwargs(small), xs, {"transform": 2, "precision": 32})):
        with CostProblem(**kw, tune=tune) as g:
            ga, ta, gb, tb = _both(g, x0)
        print(x0.size, "cost", ta[-1], tb[-1], "grad", rel_l2(gb, ga))
        assert np.all(np.isfinite(gb)) and np.all(np.isfinite(tb))
        assert np.array_equal(ta, tb)
        assert rel_l2(gb, ga) <= (1e-6 if tune and tune.get("precision") == 32 else 1e-13)


def test_noalloc_deterministic_is_bitwise(gpu_lib):
    """With smili_tune.deterministic the two entry points agree bit for bit,
    gradient included (the scratch carries the spreading's slot array), on
    two problems in a row."""
    from smili_amd import workloads
    from smili_amd.cost import CostProblem
    rng = np.random.default_rng(63)
    big, _ = workloads.single_large(seed=1)
    small, xs = random_problem(rng, 32, 28, M=400, window=True)
    for kw, x0 in ((big.kwargs, big.x0), (problem_kwargs(small), xs)):
        with CostProblem(**kw, tune={"deterministic": 1}) as g:
            ga, ta, gb, tb = _both(g, x0)
        assert np.array_equal(ta, tb) and np.array_equal(ga, gb)


def test_noalloc_rejects_problems_beyond_the_scratch(gpu_lib):
    from smili_amd._lib import SmiliError
    from smili_amd.cost import CostProblem
    rng = np.random.default_rng(62)
    # 1024^2 with full uv coverage: its two 2055 x 1031 FP64 grid windows
    # alone need 65 MiB
    p, x0 = random_problem(rng, 1024, 1024, M=100, nfcv=50, namp=0, ncp=0, nca=0, regs=("l1",))
    with CostProblem(**problem_kwargs(p)) as g:
        import torch
        x = torch.tensor(x0, dtype=torch.float64, device="cuda")
        gr = torch.empty_like(x)
        t = torch.empty(g.nterms, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        with pytest.raises(SmiliError) as e:
            g.cost_grad_device_noalloc(x.data_ptr(), gr.data_ptr(), t.data_ptr())
    assert e.value.code == -1 and "64 MiB" in str(e.value)
