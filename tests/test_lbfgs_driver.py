"""The two L-BFGS-B drivers of imaging.lbfgs.solve (SciPy's setulb called
through its reverse-communication protocol directly, as fftim2d.f90:286-351
does with lbfgsb.f, and scipy.optimize.minimize over the same routine) give
the same iterates, bit for bit, on a smooth bound-constrained objective with
active bounds: the same x, iteration and evaluation counts, objective
history and stop message, with and without bounds, at the iteration limit
and at a stop objective."""
import numpy as np
import pytest

from smili_amd.imaging import lbfgs


class Quartic:
    """f(x) = sum a (x - b)^2 / 2 + c sum (x_{i+1} - x_i)^2 + d sum x^4 /
    4: strictly convex, b < 0 in about a third of the entries so that the
    nonnegativity bound is active there."""

    def __init__(self, n, seed=0):
        rng = np.random.default_rng(seed)
        self.a = rng.uniform(0.5, 2.0, n)
        self.b = rng.normal(0.3, 0.7, n)
        self.c, self.d = 0.3, 0.05

    def cost_grad(self, x):
        r = x - self.b
        dx = np.diff(x)
        f = 0.5 * np.sum(self.a * r * r) + self.c * np.sum(dx * dx) + 0.25 * self.d * np.sum(x ** 4)
        g = self.a * r + self.d * x ** 3
        g[:-1] -= 2 * self.c * dx
        g[1:] += 2 * self.c * dx
        return float(f), g, {"cost": float(f)}


def _both(prob, x0, **kw):
    out = {}
    for d in ("setulb", "minimize"):
        out[d] = lbfgs.solve(prob, x0, verbose=False, driver=d, **kw)
    return out["setulb"], out["minimize"]


def _same(a, b):
    (xa, ia), (xb, ib) = a, b
    assert np.array_equal(xa, xb)
    assert (ia.niter, ia.nfev, ia.message, ia.success, ia.stopped) == \
           (ib.niter, ib.nfev, ib.message, ib.success, ib.stopped)
    assert ia.history == ib.history and ia.cost == ib.cost


@pytest.mark.parametrize("nonneg", [True, False])
def test_iteration_limit_same_iterates(nonneg):
    prob = Quartic(4000)
    x0 = np.full(4000, 0.5)
    a, b = _both(prob, x0, niter=12, m=10, nonneg=nonneg)
    _same(a, b)
    # the reference's loop ends after niter + 1 iterations (fftim2d.f90:286-351)
    assert a[1].niter == 13 and "ITERATIONS" in a[1].message
    if nonneg:
        assert np.all(a[0] >= 0) and np.sum(a[0] == 0) > 500


def test_convergence_same_iterates():
    prob = Quartic(500, seed=1)
    a, b = _both(prob, np.full(500, 0.2), niter=5000, m=5, factr=1e7)
    _same(a, b)
    assert a[1].success and a[1].niter < 5000


def test_stop_cost_same_iterates():
    prob = Quartic(2000, seed=2)
    x0 = np.full(2000, 1.0)
    _, full = lbfgs.solve(prob, x0, niter=200, m=10, verbose=False, driver="minimize")
    target = full.history[9]
    a, b = _both(prob, x0, niter=200, m=10, stop_cost=target)
    _same(a, b)
    assert a[1].stopped and a[1].niter == 10 and a[1].history[-1] <= target


def test_callback_sees_read_only_iterates():
    prob = Quartic(300, seed=3)
    seen = []

    def cb(it, xk, terms):
        seen.append((it, xk.flags.writeable, float(xk.sum())))

    lbfgs.solve(prob, np.full(300, 0.5), niter=5, verbose=False, driver="setulb", callback=cb)
    assert [s[0] for s in seen] == list(range(1, 7)) and not any(s[1] for s in seen)


def test_unknown_driver_is_rejected():
    with pytest.raises(ValueError):
        lbfgs.solve(Quartic(10), np.ones(10), niter=2, verbose=False, driver="fortran")


class WrongGradient(Quartic):
    """The gradient of -f: no descent direction, so the line search fails
    (lbfgsb.f's ABNORMAL_TERMINATION_IN_LNSRCH, lbfgsb.f:790-804)."""

    def cost_grad(self, x):
        f, g, t = super().cost_grad(x)
        return f, -g, t


def test_line_search_failure_is_surfaced():
    prob = WrongGradient(200, seed=5)
    a, b = _both(prob, np.full(200, 0.5), niter=50, m=5)
    _same(a, b)
    info = a[1]
    assert not info.success and "ABNORMAL" in info.message and info.niter < 50
