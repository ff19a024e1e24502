"""Host logic of the autotuner (smili_amd/autotune.py) with stand-in
handles: coordinate descent keeps the fastest setting, records rejected
settings, and refuses a setting that changes the cost."""
import pytest

from smili_amd import autotune as A
from smili_amd._lib import SmiliError


class _Fake:
    def __init__(self, tune, cost=1.0):
        self.tune, self.cost, self.closed = tune, cost, False

    def cost_grad(self, x):
        return self.cost, None, None

    def close(self):
        self.closed = True


def _time(g, x):
    t = g.tune
    return 1.0 - 0.1 * (t.get("chunk") == 2048) - 0.05 * (t.get("fft_radix") == 8) + 0.2 * (t.get("bin") == 8)


def test_coordinate_descent_finds_the_fastest_setting():
    made = []

    def make(t):
        if t.get("col_group") == 4:
            raise SmiliError(-1, "LDS budget exceeded: lower col_group or bin")
        made.append(_Fake(t))
        return made[-1]

    best, rows = A.autotune(make, None, timer=_time)
    assert best == {"chunk": 2048, "fft_radix": 8}
    assert rows[0]["tune"] == {} and rows[0]["seconds"] == 1.0
    bad = [r for r in rows if r["error"]]
    assert len(bad) == 1 and bad[0]["tune"]["col_group"] == 4 and "LDS" in bad[0]["error"]
    # every candidate once, every handle closed
    keys = [tuple(sorted(r["tune"].items())) for r in rows]
    assert len(keys) == len(set(keys)) and all(g.closed for g in made)
    assert min(r["seconds"] for r in rows if r["seconds"]) == pytest.approx(0.85)


def test_a_setting_that_changes_the_cost_is_an_error():
    def make(t):
        return _Fake(t, cost=1.0 + (1e-6 if t.get("bin") == 32 else 0.0))

    with pytest.raises(RuntimeError, match="changes the cost"):
        A.autotune(make, None, timer=_time)
    # within the FP32 tolerance it is accepted
    best, _ = A.autotune(make, None, timer=_time, base={"precision": 32}, rtol=1e-5)
    assert best["precision"] == 32


def test_imaging_rejects_unknown_tune_strings():
    from smili_amd.imaging import lbfgs
    with pytest.raises(ValueError, match="auto"):
        lbfgs.imaging(None, tune="fastest")
