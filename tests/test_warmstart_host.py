"""Host logic of the learned warm start (no GPU): seeded cases, the disk
and crescent transforms, the symmetries of the square, shard files, the
learning-rate schedule and the time-to-image bookkeeping."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from smili_amd import synth
from smili_amd.synth import UAS
from smili_amd.warmstart import data as D
from smili_amd.warmstart import gen, solve, train


def test_cases_are_seeded_and_in_range():
    a, b = D.make_case(17), D.make_case(17)
    assert a.kind == b.kind and np.array_equal(a.u, b.u) and np.array_equal(a.vis, b.vis)
    c = D.make_case(18)
    assert not (c.u.size == a.u.size and np.array_equal(c.vis, a.vis))
    kinds = {D.make_case(s).kind for s in range(40)}
    assert kinds == {"ring", "crescent", "gaussians", "disk"}
    for s in range(20):
        k = D.make_case(s)
        assert 0.3 <= k.flux <= 2.0 and k.u.size >= D.CFG["min_vis"]
        assert np.allclose(k.sigma, D.CFG["noise"] * k.flux)
        # zero-spacing flux of the analytic sky is the case's flux
        assert abs(synth.sky_vis(k.sky, [0.0], [0.0])[0] - k.flux) < 1e-12


def _raster_vis(fn_inside, half, n, u, v):
    """Fourier sum exp(+2 pi i (u x + v y)) of a unit-brightness region
    sampled on an n x n grid over [-half, half]^2 (flux-normalised)."""
    g = (np.arange(n) + 0.5) / n * 2 * half - half
    X, Y = np.meshgrid(g, g)
    m = fn_inside(X, Y).astype(float)
    m /= m.sum()
    return np.array([np.sum(m * np.exp(2j * math.pi * (uu * X + vv * Y))) for uu, vv in zip(u, v)])


def test_disk_and_crescent_transforms_match_direct_sums():
    d = 40 * UAS
    u = np.array([0.0, 3e9, 7e9, 1.2e10, -5e9])
    v = np.array([0.0, 1e9, -4e9, 2e9, 6e9])
    ref = _raster_vis(lambda X, Y: X * X + Y * Y <= (d / 2) ** 2, 25 * UAS, 1200, u, v)
    got = synth.sky_vis([synth.Disk(1.0, d)], u, v)
    assert np.abs(got - ref).max() < 2e-3
    # crescent: disk of 50 uas minus an equal-brightness disk of 30 uas
    # shifted by (6, -4) uas
    c = synth.Crescent(1.0, 50 * UAS, 30 * UAS, 6 * UAS, -4 * UAS)
    sx, sy = 6 * UAS, -4 * UAS
    inside = lambda X, Y: ((X * X + Y * Y) <= (25 * UAS) ** 2) & (((X - sx) ** 2 + (Y - sy) ** 2) > (15 * UAS) ** 2)
    ref = _raster_vis(inside, 30 * UAS, 1400, u, v)
    assert np.abs(synth.sky_vis([c], u, v) - ref).max() < 2e-3


def test_dihedral_group():
    a = np.random.default_rng(0).standard_normal((6, 6))
    outs = [D.dihedral(a, k) for k in range(8)]
    assert len({o.tobytes() for o in outs}) == 8
    # k = 1: out[i][j] = in[j][n-1-i]; k = 4: transpose
    assert np.array_equal(outs[1], np.array([[a[j, 5 - i] for j in range(6)] for i in range(6)]))
    assert np.array_equal(outs[4], a.T)
    # the same symmetry on the channels of an input stack
    x = np.stack([a, 2 * a], axis=-1)
    assert np.array_equal(D.dihedral(x, 5)[..., 1], 2 * D.dihedral(a, 5))


def test_shards_round_trip(tmp_path):
    metas = []
    ts = []
    for s in (5, 3, 9):
        ts.append(np.full((4, 4), s, np.float32))
        metas.append(dict(seed=s, kind="disk", nvis=100 + s, niter=7, nfev=8, cost=1.5, solve_s=0.1, total_s=0.2,
                          flux=1.0, dirty_peak=2.0, target_peak=3.0))
    gen._save(str(tmp_path / "shard_00.npz"), ts[:2], metas[:2])
    gen._save(str(tmp_path / "shard_01.npz"), ts[2:], metas[2:])
    _, t, m = gen.load(str(tmp_path), inputs=False)
    assert list(m["seed"]) == [3, 5, 9] and np.array_equal(t[:, 0, 0], [3, 5, 9])
    assert list(m["nvis"]) == [103, 105, 109]
    gen._save(str(tmp_path / "shard_02.npz"), ts[:1], metas[:1])
    with pytest.raises(ValueError, match="duplicate"):
        gen.load(str(tmp_path), inputs=False)


def test_shards_with_stored_inputs(tmp_path):
    """--inputs shards carry the network inputs (float16), and load returns
    them in seed order without recomputing them."""
    metas, ts, xs = [], [], []
    for s in (8, 2, 6):
        ts.append(np.full((4, 4), s, np.float32))
        xs.append(np.full((4, 4, 2), 0.5 * s, np.float32))
        metas.append(dict(seed=s, kind="ring", nvis=400, niter=100, nfev=110, cost=2.0, solve_s=0.1, total_s=0.2,
                          flux=1.0, dirty_peak=2.0, target_peak=3.0))
    gen._save(str(tmp_path / "shard_00.npz"), ts[:2], metas[:2], xs[:2])
    gen._save(str(tmp_path / "shard_01.npz"), ts[2:], metas[2:], xs[2:])
    x, t, m = gen.load(str(tmp_path))
    assert list(m["seed"]) == [2, 6, 8]
    assert x.dtype == np.float32 and np.array_equal(x[:, 0, 0, 1], [1.0, 3.0, 4.0])
    assert np.array_equal(t[:, 0, 0], [2, 6, 8])


def test_lr_schedule():
    lr = train.lr_schedule(1000, 1e-3, 100)
    assert lr.shape == (1000,) and lr.dtype == np.float32
    assert np.isclose(lr[0], 1e-5) and np.isclose(lr[99], 1e-3)
    assert np.all(np.diff(lr[100:]) <= 1e-12) and np.isclose(lr[-1], 1e-5, rtol=1e-2)


def test_time_to_image_bookkeeping():
    info = SimpleNamespace(history=[10.0, 3.0, 2.0000001, 2.0], history_t=[0.1, 0.2, 0.3, 0.4])
    assert solve._reach(info, 2.0 * (1 + 1e-6)) == (3, 0.3)
    assert solve._reach(info, 1.0) == (None, None)
    rows = [dict(reached=True, plain_iters=100, warm_iters=50, plain_s=1.0, warm_s=0.5, nrmse=1e-3, gap_prior=10.0,
                 gap_proposal=1.0, warm_start="proposal"),
            dict(reached=True, plain_iters=90, warm_iters=90, plain_s=0.9, warm_s=0.9, nrmse=2e-3, gap_prior=20.0,
                 gap_proposal=3.0, warm_start="proposal"),
            dict(reached=False, plain_iters=80, warm_iters=None, plain_s=0.8, warm_s=None, warm_wall_s=1.6,
                 nrmse=1.0, gap_prior=5.0, gap_proposal=9.0, warm_start="prior")]
    s = solve.summary(rows)
    assert s["n"] == 3 and s["reached"] == 2 and s["proposals_used"] == 2
    # the unreached row counts as a loss: 0.8 / 1.6; ratios 2, 1, 0.5
    assert np.isclose(s["median_iter_ratio"], 1.5) and np.isclose(s["median_time_ratio"], 1.0)
    assert np.isclose(s["geomean_time_ratio"], 1.0) and s["max_nrmse"] == 2e-3
    # an unreached row never counts as a gain
    assert solve.time_ratio(dict(reached=False, plain_s=5.0, warm_wall_s=1.0)) == 1.0
