"""Writes the golden fixtures of tests/golden/ (run in the build container).

The reference ships no tests, golden vectors or fixtures with expected
outputs (its only data are input tables and images under docs/examples), so
every fixture here is produced by the project's FP64 oracle:

* cost_<name>.npz -- small calc_cost cases (inputs + exact-DFT cost,
  gradient and the 15 terms);
* solve_small_fcv.npz -- the 64x64 full-visibility workload (fcv chi-square +
  L1 (1) + TSV (1e4), non-negative, L-BFGS-B m=10, 1000 iterations from a
  flat 20 uas Gaussian) solved with the exact-DFT oracle cost and SciPy's
  L-BFGS-B (the algorithm of the reference's src/lbfgsb/lbfgsb.f v3.0): the
  final image the GPU imager must reproduce.

Usage:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
from scipy.optimize import minimize

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import smili_oracle as O  # noqa: E402
from tests.helpers import problem_kwargs, random_problem  # noqa: E402

CASES = {
    "all_terms_24x20": dict(seed=101, shape=(24, 20), kw=dict(M=300)),
    "window_16x16": dict(seed=102, shape=(16, 16), kw=dict(M=200, window=True)),
    "moments_12x14": dict(seed=103, shape=(12, 14), kw=dict(M=120, extra_regs=("tfd", "cen", "sm"))),
    "odd_13x11": dict(seed=104, shape=(13, 11), kw=dict(M=150, regs=("l1", "tv"))),
}


def write_cost_cases():
    for name, c in CASES.items():
        rng = np.random.default_rng(c["seed"])
        p, x = random_problem(rng, *c["shape"], **c["kw"])
        cost, grad, terms = O.calc_cost(x, p)
        kw = problem_kwargs(p)
        arrays = {"in_" + k: np.asarray(v) for k, v in kw.items()}
        np.savez_compressed(os.path.join(HERE, "cost_%s.npz" % name), x=x, cost=cost, grad=grad,
                            terms=O.terms_vector(terms), **arrays)
        print(name, cost)


def write_small_fcv_solve(niter=1000, m=10):
    from smili_amd import workloads
    prep, extra = workloads.small_fcv()
    p = O.Problem(**prep.kwargs)
    nfev = [0]

    def fg(x):
        nfev[0] += 1
        c, g, _ = O.calc_cost(x, p)
        return c, g

    res = minimize(fg, prep.x0, jac=True, method="L-BFGS-B", bounds=[(0, None)] * prep.x0.size,
                   options=dict(maxcor=m, ftol=10 * np.finfo(float).eps, gtol=0.0, maxiter=niter + 1,
                                maxfun=10 ** 9, maxls=20))
    img = np.zeros((prep.Ny, prep.Nx))
    img[prep.yidx - 1, prep.xidx - 1] = res.x / prep.inormfactr
    np.savez_compressed(os.path.join(HERE, "solve_small_fcv.npz"), image=img, cost=res.fun,
                        nit=res.nit, nfev=nfev[0], x0=prep.x0, inormfactr=prep.inormfactr)
    print("small_fcv solve: nit", res.nit, "nfev", nfev[0], "cost", res.fun, res.message)


if __name__ == "__main__":
    write_cost_cases()
    write_small_fcv_solve()
