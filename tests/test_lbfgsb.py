"""The native L-BFGS-B (smili_amd.lbfgsb, libsmili_lbfgsb.so), which stands
in for the reference's src/lbfgsb/lbfgsb.f:

* against SciPy's implementation of the same algorithm (v3.0, driven through
  its setulb protocol like fftim2d.f90:286-351): the same iterates to
  rounding for the first iterations of a bound-constrained quartic with
  active bounds, of a convex quadratic and of the Rosenbrock function
  (rounding then grows on Rosenbrock's curved valley; tolerances below),
  and the same minimum;
* closed forms: a separable quadratic in a box ends at clip(b / a, l, u);
  Rosenbrock at all ones; an upper-bounded problem at its bound;
* the protocol: NEW_X after every iteration, the iteration limit of the
  imager's loop (niter + 1 iterations), a failed line search surfaced as
  ABNORMAL_TERMINATION_IN_LNSRCH, invalid set-ups rejected;
* batched stepping = one problem at a time, bit for bit; threaded O(n m)
  work (n >= 65536) = one thread up to the reductions' rounding."""
import numpy as np
import pytest

from smili_amd import lbfgsb as LB
from smili_amd.imaging import lbfgs
from tests.test_lbfgs_driver import Quartic, WrongGradient


def _scipy_iterates(fg, x0, nonneg, m, factr, pgtol, maxiter):
    xs = []
    out = lbfgs._setulb_loop(fg, x0, nonneg, m, factr, pgtol, maxiter, 20, lambda x: xs.append(np.array(x)))
    assert out is not None
    return out, xs


def _native_iterates(fg, x0, nonneg, m, factr, pgtol, maxiter, nthreads=1):
    xs = []
    out = LB.minimize(fg, x0, m=m, lower=0.0 if nonneg else None, factr=factr, pgtol=pgtol, maxiter=maxiter,
                      callback=lambda x: xs.append(np.array(x)), nthreads=nthreads)
    return out, xs


def _fg(prob):
    return lambda x: prob.cost_grad(x)[:2]


def test_quartic_iterates_match_scipy():
    prob = Quartic(3000, seed=4)
    x0 = np.full(3000, 0.5)
    (xs_, nit_s, msg_s, ok_s), it_s = _scipy_iterates(_fg(prob), x0, True, 10, 1e1, 0.0, 40)
    (xn, nit_n, msg_n, ok_n), it_n = _native_iterates(_fg(prob), x0, True, 10, 1e1, 0.0, 40)
    assert nit_s == nit_n and ok_s and ok_n              # both stop by the f test, at the same iteration
    for k in range(nit_n):
        assert np.max(np.abs(it_n[k] - it_s[k])) <= 1e-14 * np.max(np.abs(it_s[k])), k
    assert np.sum(xn == 0.0) > 500                      # the bound is active
    assert np.array_equal(xn == 0.0, xs_ == 0.0)
    assert prob.cost_grad(xn)[0] == pytest.approx(prob.cost_grad(xs_)[0], rel=1e-13)


def test_quadratic_iterates_match_scipy_to_convergence():
    rng = np.random.default_rng(0)
    n = 60
    A = rng.standard_normal((n, n))
    H = A @ A.T / n + 0.1 * np.eye(n)
    b = rng.standard_normal(n)

    def fg(x):
        return 0.5 * x @ H @ x - b @ x, H @ x - b

    x0 = rng.uniform(0, 1, n)
    (xs_, nit_s, _, ok_s), it_s = _scipy_iterates(fg, x0, True, 5, 1e1, 0.0, 200)
    (xn, nit_n, msg, ok_n), it_n = _native_iterates(fg, x0, True, 5, 1e1, 0.0, 200)
    assert ok_s and ok_n and "CONVERGENCE" in msg and abs(nit_s - nit_n) <= 2
    for k in range(min(nit_s, nit_n)):
        assert np.max(np.abs(it_n[k] - it_s[k])) <= 1e-14 * np.max(np.abs(it_s[k])) + 1e-15, k


def _rosen(x):
    f = np.sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] = -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
    g[1:] += 200 * (x[1:] - x[:-1] ** 2)
    return f, g


def test_rosenbrock_matches_scipy_then_reaches_ones():
    x0 = np.random.default_rng(1).uniform(-1, 1, 20)
    (_, _, _, ok_s), it_s = _scipy_iterates(_rosen, x0, False, 10, 1e7, 1e-5, 500)
    (xn, nit, msg, ok), it_n = _native_iterates(_rosen, x0, False, 10, 1e1, 1e-10, 2000)
    # rounding differences grow by ~5x per iteration along the valley
    for k in range(8):
        assert np.max(np.abs(it_n[k] - it_s[k])) <= 1e-15, k
    for k in range(8, 12):
        assert np.max(np.abs(it_n[k] - it_s[k])) <= 1e-12, k
    assert ok and np.max(np.abs(xn - 1.0)) <= 1e-6


def test_box_quadratic_closed_form():
    rng = np.random.default_rng(2)
    n = 500
    a = rng.uniform(0.5, 3.0, n)
    b = rng.normal(0.0, 2.0, n)
    lo = rng.uniform(-1.0, 0.0, n)
    up = lo + rng.uniform(0.2, 2.0, n)
    lo[:50] = -np.inf                                   # upper bound only
    up[50:100] = np.inf                                 # lower bound only
    lo[100:120] = -np.inf
    up[100:120] = np.inf                                # free
    lo[120:130] = up[120:130] = 0.25                    # fixed (l = u)

    def fg(x):
        return float(np.sum(0.5 * a * x * x - b * x)), a * x - b

    x, nit, msg, ok = LB.minimize(fg, np.zeros(n), m=5, lower=lo, upper=up, factr=0.0, pgtol=1e-11, maxiter=500)
    assert ok, msg
    # f stagnates at rounding level first: x is exact to ~sqrt(eps)
    np.testing.assert_allclose(x, np.clip(b / a, lo, up), rtol=0, atol=1e-6)
    at = (x == lo) | (x == up)
    assert np.array_equal(at, (b / a <= lo) | (b / a >= up))   # the active set exactly
    assert np.all(x[120:130] == 0.25)


def test_protocol_and_iteration_limit():
    prob = Quartic(400, seed=6)
    x = np.full(400, -1.0)                              # projected at START
    with LB.Lbfgsb(400, m=5, lower=0.0) as opt:
        task = opt.step(x)
        assert task == LB.FG and np.all(x == 0.0)
        tasks, nfev = [], 0
        while len(tasks) < 30 and task in (LB.FG, LB.NEW_X):
            if task == LB.FG:
                f, g = _fg(prob)(x)
                nfev += 1
                task = opt.step(x, f, g)
            else:
                tasks.append(task)
                task = opt.step(x)
        it, nf, _ = opt.info()
        assert it == len(tasks) and nf == nfev
    # the imager's loop: niter + 1 iterations (fftim2d.f90:286-351)
    x, info = lbfgs.solve(prob, np.full(400, 0.5), niter=12, m=5, verbose=False, driver="native")
    assert info.niter == 13 and "ITERATIONS" in info.message and not info.success


def test_line_search_failure_is_abnormal():
    prob = WrongGradient(200, seed=5)
    _, info = lbfgs.solve(prob, np.full(200, 0.5), niter=50, m=5, verbose=False, driver="native")
    assert not info.success and "ABNORMAL" in info.message and info.niter < 50


def test_invalid_setups_are_rejected():
    with pytest.raises(ValueError):
        LB.Lbfgsb(0, m=5)
    with pytest.raises(ValueError):
        LB.Lbfgsb(10, m=0)
    with pytest.raises(ValueError):
        LB.Lbfgsb(10, m=5, lower=np.ones(10), upper=np.zeros(10))
    with LB.Lbfgsb(10, m=5) as opt:
        with pytest.raises(ValueError):
            opt.step(np.zeros(9))


def test_batch_equals_single():
    probs = [Quartic(n, seed=s) for n, s in ((300, 1), (500, 2), (200, 3))]
    x0s = [np.full(p.a.size, 0.4) for p in probs]
    singles = [LB.minimize(_fg(p), x0, m=5, lower=0.0, factr=1e1, pgtol=0.0, maxiter=25)[0]
               for p, x0 in zip(probs, x0s)]
    xs = [x0.copy() for x0 in x0s]
    batch = LB.BatchLbfgsb([p.a.size for p in probs], m=5, lower=0.0, factr=1e1, pgtol=0.0, nthreads=3)
    tasks = batch.step(xs)
    iters = np.zeros(len(probs), int)
    fs, gs = np.zeros(len(probs)), [np.zeros(p.a.size) for p in probs]
    while True:
        for k, p in enumerate(probs):
            if tasks[k] == LB.FG:
                fs[k], gs[k] = _fg(p)(xs[k])
            elif tasks[k] == LB.NEW_X:
                iters[k] += 1
        # a problem at its 25th NEW_X is not stepped again (the loop of
        # minimize stops there)
        active = [t == LB.FG or (t == LB.NEW_X and iters[k] < 25) for k, t in enumerate(tasks)]
        if not any(active):
            break
        tasks = batch.step(xs, fs, gs, active)
    batch.close()
    for x, s in zip(xs, singles):
        assert np.array_equal(x, s)


def test_threads_match_one_thread():
    n = 1 << 17
    prob = Quartic(n, seed=7)
    x0 = np.full(n, 0.5)
    x1 = LB.minimize(_fg(prob), x0, m=5, lower=0.0, factr=1e1, pgtol=0.0, maxiter=15, nthreads=1)[0]
    x8 = LB.minimize(_fg(prob), x0, m=5, lower=0.0, factr=1e1, pgtol=0.0, maxiter=15, nthreads=8)[0]
    assert np.max(np.abs(x8 - x1)) <= 1e-10 * np.max(np.abs(x1))


def test_library_exports_every_declared_symbol():
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    txt = open(os.path.join(root, "include", "smili_lbfgsb.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(smili_\w+)\s*\(", txt)))
    assert len(names) == 5
    L = LB.lib()
    for n in names:
        assert hasattr(L, n), n
