"""CPU: the oracle reproduces the committed golden fixtures, and the
product's set-up code (smili_amd.imaging.prep) agrees with the oracle's
restatement of lbfgs.py / fftim2d.f90 / tools.py."""
import glob
import os

import numpy as np
import pytest

from oracle import smili_oracle as O
from smili_amd.imaging import prep as P

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_case(path):
    z = np.load(path, allow_pickle=False)
    kw = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    for k, v in list(kw.items()):
        if v.ndim == 0:
            kw[k] = int(v) if k in ("Nx", "Ny") else float(v)
    return O.Problem(**kw), z["x"], float(z["cost"]), z["grad"], z["terms"]


CASES = sorted(glob.glob(os.path.join(HERE, "cost_*.npz")))


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(c) for c in CASES])
def test_oracle_reproduces_golden_costs(path):
    p, x, cost, grad, terms = load_case(path)
    c, g, t = O.calc_cost(x, p)
    assert abs(c - cost) <= 1e-12 * abs(cost)
    assert np.allclose(g, grad, rtol=1e-10, atol=1e-12 * np.abs(grad).max())
    assert np.allclose(O.terms_vector(t), terms, rtol=1e-10, atol=1e-300)


def test_get_uvlist_product_matches_oracle():
    rng = np.random.default_rng(1)
    n = 40
    fcv = dict(u=rng.normal(size=n), v=rng.normal(size=n))
    amp = dict(u=np.r_[fcv["u"][:10], rng.normal(size=5)], v=np.r_[-fcv["v"][:10], rng.normal(size=5)])
    cp = {k: rng.normal(size=7) for k in ("u12", "v12", "u23", "v23")}
    cp["u31"], cp["v31"] = -(cp["u12"] + cp["u23"]), -(cp["v12"] + cp["v23"])
    ca = {k: rng.normal(size=6) for k in ("u1", "v1", "u2", "v2", "u3", "v3", "u4", "v4")}
    a = P.get_uvlist(fcv, amp, cp, ca)
    b = O.get_uvlist((fcv["u"], fcv["v"]), (amp["u"], amp["v"]),
                     ((cp["u12"], cp["v12"]), (cp["u23"], cp["v23"]), (cp["u31"], cp["v31"])),
                     tuple((ca["u%d" % k], ca["v%d" % k]) for k in range(1, 5)))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    u, v, ifcv = a[0], a[1], a[2]
    assert np.all(v >= 0)
    # indices reconstruct the original coordinates (tools.py:20-27, 104-108)
    s = np.sign(ifcv)
    assert np.allclose(u[np.abs(ifcv) - 1] * s, fcv["u"]) and np.allclose(v[np.abs(ifcv) - 1] * s, fcv["v"])
    # amp entries at (u, -v) of fcv entries map to distinct points with sign
    assert len(np.unique(u + 1j * v)) == u.size


def test_weight_initialisers_product_matches_oracle():
    rng = np.random.default_rng(2)
    prior = rng.uniform(0.1, 1.0, 30)
    xi = np.tile(np.arange(1, 7), 5)
    yi = np.repeat(np.arange(1, 6), 6)
    assert np.allclose(P.init_l1reg(prior), O.init_l1reg(prior), rtol=1e-14)
    for flat in (True, False):
        # reweighted TV divides by tv_e, which is 0 at the last pixel (no
        # forward neighbours): the reference's weights are then NaN as well
        # (calc_tv_wgt, image.f90:437-441)
        assert np.allclose(P.init_tvreg(xi, yi, 6, 5, flat, prior), O.init_tvreg(xi, yi, 6, 5, flat, prior),
                           rtol=1e-13, equal_nan=True)
        assert np.allclose(P.init_tsvreg(xi, yi, 6, 5, flat, prior), O.init_tsvreg(xi, yi, 6, 5, flat, prior),
                           rtol=1e-13)
    la, wa = P.init_klreg(3.0, prior)
    lb, wb = O.init_klreg(3.0, prior)
    assert abs(la - lb) < 1e-14 * la and np.allclose(wa, wb)
    assert np.allclose(P.init_gsreg(prior), O.init_gsreg(prior))
    assert P.init_tfdreg(1.0, 2.0, 0.01) == O.init_tfdreg(1.0, 2.0, 0.01)


def test_prepare_matches_oracle_setup():
    """Weights normalisation (fftim2d.f90:178-213) and the phase shift of the
    complex visibilities (fftim2d.f90:224-241)."""
    from smili_amd import workloads
    prep, extra = workloads.small_fcv()
    kw = prep.kwargs
    vis = extra["vis"]
    wf, wa, wc, wq = O.normalise_weights(len(vis["u"]), 1, 0, 1, 0, 1, 0, 1)
    assert abs(kw["wfcv"] - wf) < 1e-18
    u, v, ifcv, *_ = O.get_uvlist((vis["u"], vis["v"]))
    up = u * 2 * np.pi * np.deg2rad(-2 / 3.6e9)
    vp = v * 2 * np.pi * np.deg2rad(2 / 3.6e9)
    V = vis["amp"] * np.exp(1j * np.deg2rad(vis["phase"])) * prep.inormfactr
    Vs = O.phashift_r2c(V, ifcv, up, vp, 32.5, 32.5, 64, 64)
    assert np.allclose(kw["Vfcv"], Vs, rtol=1e-12, atol=1e-12 * np.abs(Vs).max())
    assert np.allclose(kw["u"], up) and np.allclose(kw["v"], vp)


def test_image_container_grid_and_nrmse():
    from smili_amd.imaging.image import Image, nrmse
    im = Image(8, 6, 2.0, angunit="uas")
    xg, yg = im.get_xygrid()
    assert np.isclose(xg[0], 2.0 * (0 - 4.5 + 1) * -1.0)
    assert im.header["nxref"] == 4.5 and im.header["dx"] < 0
    a = np.ones((6, 8))
    assert nrmse(a, a) == 0.0
    assert np.isclose(nrmse(2 * a, a), 1.0)
