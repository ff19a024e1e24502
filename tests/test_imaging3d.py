"""Host set-up of the dynamic imager (smili/imaging/lbfgs3d.py:256-738 and
tools.get_uvlist_loop, smili/uvdata/uvtable/tools.py:255-336) on CPU: the
concatenated uv list reproduces every table row's (u, v) through its signed
index, frames own consecutive uv ranges, and the weights follow the
reference's initialisers."""
import numpy as np
import pytest

from smili_amd import synth
from smili_amd.imaging.prep3d import get_uvlist_loop, prepare3d

UAS = synth.UAS


def _movie_tables(rng, Nt=4, n_per=(30, 0, 45, 20), with_closures=True):
    vis = {k: [] for k in ("u", "v", "amp", "phase", "sigma", "frmidx")}
    bs = {k: [] for k in ("u12", "v12", "u23", "v23", "u31", "v31", "amp", "phase", "sigma", "frmidx")}
    for z in range(Nt):
        n = n_per[z]
        if n == 0:
            continue
        u = rng.uniform(-5e9, 5e9, n)
        v = rng.uniform(-5e9, 5e9, n)
        u[::7] = u[0]                      # repeated points inside a frame
        v[::7] = v[0]
        for k, a in (("u", u), ("v", v), ("amp", rng.uniform(0.1, 1, n)), ("phase", rng.uniform(-180, 180, n)),
                     ("sigma", np.full(n, 0.01)), ("frmidx", np.full(n, z))):
            vis[k].append(a)
        if with_closures:
            m = n // 3
            a, b, c = u[:m], u[m:2 * m], u[2 * m:3 * m]
            va, vb, vc = v[:m], v[m:2 * m], v[2 * m:3 * m]
            for k, arr in (("u12", a), ("v12", va), ("u23", b), ("v23", vb), ("u31", c), ("v31", vc),
                           ("amp", np.ones(m)), ("phase", rng.uniform(-30, 30, m)), ("sigma", np.full(m, 0.05)),
                           ("frmidx", np.full(m, z))):
                bs[k].append(arr)
    return ({k: np.concatenate(a) for k, a in vis.items()},
            {k: np.concatenate(a) for k, a in bs.items()} if with_closures else None)


def test_uvlist_loop_indices_reproduce_rows():
    rng = np.random.default_rng(0)
    vis, bs = _movie_tables(rng)
    u, v, ifcv, iamp, icp, ica, Nuvs, order = get_uvlist_loop(4, vis, None, bs, None)
    assert Nuvs[1] == 0 and Nuvs.sum() == u.size
    start = np.concatenate([[0], np.cumsum(Nuvs)])
    rows = order[0]
    # every visibility row maps to its own (u, v) (or its conjugate point)
    j = np.abs(ifcv) - 1
    s = np.sign(ifcv)
    np.testing.assert_array_equal(u[j] * s, vis["u"][rows])
    np.testing.assert_array_equal(v[j] * s, vis["v"][rows])
    # ... inside its frame's range of the concatenated list
    fr = vis["frmidx"][rows]
    assert np.all(j >= start[fr]) and np.all(j < start[fr + 1])
    # the uv list holds v >= 0 only and no duplicates inside a frame
    assert np.all(v >= 0)
    for z in range(4):
        seg = u[start[z]:start[z + 1]] + 1j * v[start[z]:start[z + 1]]
        assert np.unique(seg).size == seg.size
    # closure phases: three signed indices per row, the same frame
    brow = order[2]
    for k, (cu, cv) in enumerate((("u12", "v12"), ("u23", "v23"), ("u31", "v31"))):
        jj = np.abs(icp[k]) - 1
        ss = np.sign(icp[k])
        np.testing.assert_array_equal(u[jj] * ss, bs[cu][brow])
        np.testing.assert_array_equal(v[jj] * ss, bs[cv][brow])


def test_prepare3d_layout_and_weights():
    rng = np.random.default_rng(1)
    vis, bs = _movie_tables(rng)
    Nt, N = 4, 16
    frames = rng.uniform(0.1, 1.0, (Nt, N, N))
    dx = -2 * UAS
    pm = prepare3d(frames, dx, 2 * UAS, vistable=vis, bstable=bs, l1_lambda=1.0, tsv_lambda=10.0,
                   rt_lambda=2.0, ri_lambda=0.5, lc_lambda=1.0, lc_array=[1.0, 1.1, 0.9, 1.0],
                   lc_normalize="both")
    kw = pm.kwargs
    Npix = N * N
    assert pm.x0.size == Nt * Npix
    np.testing.assert_allclose(pm.x0.reshape(Nt, N, N), frames * pm.inormfactr)
    assert pm.inormfactr == pytest.approx(1.0 / frames.max())
    # every 7th point of a frame repeats its first one
    np.testing.assert_array_equal(kw["Nuvs"], [30 - 4, 0, 45 - 6, 20 - 2])
    assert kw["Nuvs"].sum() == kw["u"].size
    # flat Rt / Ri priors: w = 1 / (Nt Npix (1 / Npix)) (lbfgs3d.py:566-580)
    np.testing.assert_allclose(kw["rt_wgt"], 1.0 / Nt)
    np.testing.assert_allclose(kw["ri_wgt"], 1.0 / Nt)
    # light curve (lbfgs3d.py:538-550): targets scaled, lambda / (err tgt)^2
    lc = np.array([1.0, 1.1, 0.9, 1.0]) * pm.inormfactr
    np.testing.assert_allclose(kw["lc_tgtfd"], lc)
    np.testing.assert_allclose(kw["lc_l"], 1.0 / (0.01 * lc) ** 2)
    assert kw["lc_nidx"] == 3
    # tfd off with lc_normalize both: tfd target = mean(lc_array) * inormfactr
    assert kw["tfd_l"] == -1.0 and kw["tfd_tgtfd"] == pytest.approx(np.mean([1.0, 1.1, 0.9, 1.0]) * pm.inormfactr)
    # data weights normalised over all data (fftim3d.f90:225-259)
    nf, nc = len(kw["uvidxfcv"]), kw["uvidxcp"].shape[1]
    assert kw["wfcv"] * 2 * nf + kw["wcp"] * nc == pytest.approx(1.0)


def test_prepare3d_without_lc_array_and_flat_l1():
    """Two places where lbfgs3d.py fails are defined here: tfd off without
    lc_array (len(None), lbfgs3d.py:515) and l1 without a prior
    (lbfgs3d.py:356-359 leaves l1_wgt unset): the flat prior of the static
    imager is used."""
    rng = np.random.default_rng(2)
    vis, _ = _movie_tables(rng, with_closures=False)
    frames = np.full((4, 8, 8), 0.01)
    pm = prepare3d(frames, -2 * UAS, 2 * UAS, vistable=vis, l1_lambda=1.0)
    kw = pm.kwargs
    assert kw["lc_nidx"] == -1 and np.all(kw["lc_l"] == -1.0) and np.all(kw["lc_tgtfd"] == 1.0)
    w = kw["l1_wgt"]
    assert np.allclose(w, w[0])


def test_unsorted_frames_keep_data_with_their_points():
    """Rows of a table that is not ordered by frame still pair each datum
    with its own uv point (data arrays follow the index order)."""
    rng = np.random.default_rng(3)
    vis, _ = _movie_tables(rng, with_closures=False)
    perm = rng.permutation(vis["u"].size)
    shuf = {k: a[perm] for k, a in vis.items()}
    frames = np.full((4, 8, 8), 0.01)
    a = prepare3d(frames, -2 * UAS, 2 * UAS, vistable=vis, l1_lambda=1.0).kwargs
    b = prepare3d(frames, -2 * UAS, 2 * UAS, vistable=shuf, l1_lambda=1.0).kwargs
    np.testing.assert_array_equal(a["u"], b["u"])
    key = lambda kw: np.sort(kw["uvidxfcv"] * 1e6 + np.round(np.abs(kw["Vfcv"]), 6))
    np.testing.assert_allclose(key(a), key(b))


def test_warmstart_rejects_off_centre_or_flipped_movies():
    """imaging3d(warmstart=...) bins the network's proposals onto a grid
    centred on the reference pixel with RA to the left: other geometries
    and closure-only imaging (no visibilities to propose from) are rejected
    before anything is evaluated."""
    import pytest
    from smili_amd.imaging.image import Image, Movie
    from smili_amd.imaging.lbfgs3d import imaging3d
    vis = {"u": np.zeros(3), "v": np.zeros(3), "amp": np.ones(3), "phase": np.zeros(3),
           "sigma": np.ones(3), "frmidx": np.zeros(3, int)}
    net = object()
    ok = Movie([Image(16, 16, 2.0) for _ in range(2)])
    assert ok.images[0].header["dx"] < 0
    with pytest.raises(ValueError, match="vistable"):
        imaging3d(ok, warmstart=net)
    off = Movie([Image(16, 16, 2.0, nxref=5.0) for _ in range(2)])
    with pytest.raises(ValueError, match="reference pixel"):
        imaging3d(off, vistable=vis, warmstart=net)
    flipped = Movie([Image(16, 16, 2.0) for _ in range(2)])
    for im in flipped.images:
        im.header["dx"] = abs(im.header["dx"])
    with pytest.raises(ValueError, match="dx < 0"):
        imaging3d(flipped, vistable=vis, warmstart=net)
