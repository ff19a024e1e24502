"""The per-frame warm start of a movie: binning a network image onto a
coarser movie grid (host), the per-frame network inputs (GPU transforms)."""
import numpy as np
import pytest

from smili_amd.warmstart import movie as WM


def test_bin_to_grid_keeps_flux_and_centre():
    rng = np.random.default_rng(0)
    img = np.kron(rng.uniform(size=(4, 4)), np.ones((2, 2)))     # constant on each 2 x 2 block
    out = WM.bin_to_grid(img, 8, 2)
    assert out.shape == (8, 8) and np.isclose(out.sum(), img.sum())
    # block (b, c) of 2x2 network pixels lands on movie pixel (2 + b, 2 + c)
    assert np.isclose(out[2, 2], img[0:2, 0:2].sum()) and np.isclose(out[5, 5], img[6:8, 6:8].sum())
    assert out[:2].sum() == 0 and out[:, 6:].sum() == 0
    # the reference pixels ((n + 1) / 2, 1-based) coincide: an image constant
    # on the blocks keeps its centroid
    yy, xx = np.mgrid[0:8, 0:8]
    c_in = ((xx - 3.5) * 1.0 * img).sum() / img.sum()          # network pixel = 1
    c_out = ((xx - 3.5) * 2.0 * out).sum() / out.sum()         # movie pixel = 2
    assert np.isclose(c_in, c_out)
    with pytest.raises(ValueError):
        WM.bin_to_grid(img, 7, 2)


@pytest.mark.gpu
def test_frame_inputs_per_frame(gpu_lib):
    from smili_amd.warmstart import data as D
    cfg = dict(D.CFG, n=16)
    rng = np.random.default_rng(1)
    u = rng.uniform(-5e9, 5e9, 60)
    v = rng.uniform(0, 5e9, 60)
    vis = dict(u=u, v=v, amp=np.ones(60), phase=np.zeros(60), sigma=np.ones(60),
               frmidx=np.repeat([0, 2], 30))
    x = WM.frame_inputs(vis, 3, cfg)
    assert x.shape == (3, 16, 16, 2)
    assert np.all(x[1] == 0) and x[0, ..., 0].max() == pytest.approx(1.0)


@pytest.mark.gpu
def test_movie_warm_start_reaches_plain_objective_sooner(gpu_lib):
    """A 16-frame movie of the bench's kind: the per-frame proposal's joint
    objective is below the prior frames', the warm-started solve reaches the
    plain solve's final objective within 1e-6 relative in fewer iterations,
    and its movie is as close to the true frames as the plain one (within
    20%).  The two movies themselves differ by a few percent: with 2e3-8e3
    visibilities per 128^2 frame the objective is flat along many pixel
    combinations, so movies 1e-6 apart in objective need not coincide."""
    from smili_amd import workloads
    from smili_amd.warmstart.unet import UNet
    prep, meta = workloads.dynamic_movie(n_frames=16, n_vis=75_000)
    net, _ = UNet.from_file(batch=16)
    try:
        r = WM.time_to_image_movie(net, prep, meta["vis"], 2.0, truth=meta["frames"], track=True)
    finally:
        net.close()
    print({k: r[k] for k in ("gap_prior", "gap_proposal", "plain_iters", "warm_iters", "plain_s", "warm_s",
                             "nrmse", "nrmse_plain_at_target", "truth_err_plain", "truth_err_warm",
                             "plain_curve", "warm_curve")})
    assert r["start"] == "proposal" and r["gap_proposal"] < r["gap_prior"]
    assert r["reached"] and r["warm_iters"] < r["plain_iters"]
    assert r["truth_err_warm"] <= 1.2 * r["truth_err_plain"]
    for k in ("plain_curve", "warm_curve"):
        assert None not in r[k] and r[k] == sorted(r[k])
    assert all(w <= p for w, p in zip(r["warm_curve"], r["plain_curve"]))
    assert r["nrmse"] <= 0.1 and r["nrmse_plain_at_target"] < r["nrmse"]


def test_propose_movie_bins_normalises_and_keeps_prior_of_empty_frames(monkeypatch):
    """Host logic of the per-frame proposal with a stub network and a
    quadratic objective f(x) = |x - t|^2 / 2: frames with inputs get their
    binned network image at the prior frame's flux, a frame without
    visibilities keeps its prior frame, and the common scale is the exact
    minimiser of f along the stacked proposal (the secant is exact on a
    quadratic)."""
    from types import SimpleNamespace
    from smili_amd.warmstart import data as D
    rng = np.random.default_rng(4)
    nz, n, n_m = 3, 8, 6
    cfg = dict(D.CFG, n=n, pix_uas=1.0)
    x_in = rng.uniform(size=(nz, n, n, 2)).astype(np.float32)
    x_in[1] = 0.0                                                 # frame 1: no visibilities
    shapes = rng.uniform(size=(nz, n, n)).astype(np.float32)
    monkeypatch.setattr(WM, "frame_inputs", lambda vis, nz_, cfg_: x_in)
    net = SimpleNamespace(forward=lambda x: shapes)
    yy, xx = np.mgrid[1:n_m + 1, 1:n_m + 1]
    prior = rng.uniform(0.5, 1.5, (nz, n_m * n_m))
    prep = SimpleNamespace(Nt=nz, Nx=n_m, Ny=n_m, xidx=xx.ravel(), yidx=yy.ravel(), x0=prior.ravel())
    target = rng.uniform(size=nz * n_m * n_m)
    prob = SimpleNamespace(cost_grad=lambda x: (0.5 * np.sum((x - target) ** 2), x - target, {}))
    x0, tm = WM.propose_movie(net, None, prep, prob, 2.0, cfg)
    assert tm["frames_proposed"] == 2
    s = np.empty((nz, n_m * n_m))
    for z in range(nz):
        img = WM.bin_to_grid(shapes[z].astype(np.float64), n_m, 2).ravel()
        s[z] = prior[z] if z == 1 else img * prior[z].sum() / img.sum()
    a = float(s.ravel() @ target / (s.ravel() @ s.ravel()))
    np.testing.assert_allclose(x0, a * s.ravel(), rtol=1e-10)
    # no frame with a proposal: no start
    x_in[:] = 0.0
    x0, tm = WM.propose_movie(net, None, prep, prob, 2.0, cfg)
    assert x0 is None and tm["frames_proposed"] == 0
    with pytest.raises(ValueError):
        WM.propose_movie(net, None, prep, prob, 1.5, cfg)


@pytest.mark.gpu
def test_imaging3d_warmstart_keyword(gpu_lib):
    """imaging3d(..., warmstart=net) on a 16-frame movie starts from the
    per-frame proposal and converges to the plain solve's objective (1e-4
    relative; both run to L-BFGS-B's own stop)."""
    from smili_amd import workloads
    from smili_amd.imaging import lbfgs3d
    from smili_amd.imaging.image import Movie
    from smili_amd.warmstart.unet import UNet
    prep, meta = workloads.dynamic_movie(n_frames=16, n_vis=75_000)
    frames0 = prep.x0.reshape(16, 128, 128) / prep.inormfactr
    regs = dict(l1_lambda=1.0, tsv_lambda=1e4, rt_lambda=1.0)
    movie = Movie.from_frames(frames0, dx=2.0, angunit="uas")
    _, ip = lbfgs3d.imaging3d(movie, vistable=meta["vis"], verbose=False, return_info=True, **regs)
    net, _ = UNet.from_file(batch=16)
    try:
        out, iw = lbfgs3d.imaging3d(movie, vistable=meta["vis"], verbose=False, return_info=True,
                                    warmstart=net, **regs)
    finally:
        net.close()
    print("plain", ip.niter, ip.cost, "warm", iw.start, iw.niter, iw.cost)
    assert ip.start == "prior" and iw.start == "proposal"
    assert iw.cost <= ip.cost * (1 + 1e-4)
    assert out.frames().shape == (16, 128, 128)
