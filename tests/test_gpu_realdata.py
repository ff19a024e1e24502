"""GPU cost and imager on the reference's example data (VLBA 3C273 at
43 GHz: visibilities, closure phases and log closure amplitudes of
tests/golden/uvdata_3c273.npz) at the reference's example geometry: 256x256,
20 uas pixels, reference pixel (40, 216) (docs/examples/uvdata/
3C273_image.fits header).  FP64 NUFFT against the FP64 ES-kernel oracle:
cost rel <= 1e-11, gradient rel-l2 <= 1e-10; 20 iterations of imaging()
against the same L-BFGS-B driven by the oracle: NRMSE <= 1e-6."""
import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import rel_l2
from tests.test_uvdata_realdata import load_tables

pytestmark = pytest.mark.gpu
ES = dict(fwd=O.es_nufft_type2, adj=O.es_nufft_type1)


def _setup():
    from smili_amd import synth
    from smili_amd.imaging.image import Image
    img = Image(256, 256, 20.0, angunit="uas", nxref=40.0, nyref=216.0)
    h = img.header
    img.data[0, 0] = synth.sky_image([synth.Gaussian(5.0, 150 * synth.UAS)], 256, 256,
                                     np.deg2rad(h["dx"]), np.deg2rad(h["dy"]), h["nxref"], h["nyref"])
    return img


REGS = dict(l1_lambda=1.0, tsv_lambda=1.0, kl_lambda=0.1)


def test_cost_on_real_data_matches_oracle(gpu_lib):
    from smili_amd.cost import CostProblem
    from smili_amd.imaging.prep import prepare
    vis, bs, ca = load_tables()
    img = _setup()
    h = img.header
    p = prepare(img.data[0, 0], np.deg2rad(h["dx"]), np.deg2rad(h["dy"]), h["nxref"], h["nyref"],
                vistable=vis, bstable=bs, catable=ca, **REGS)
    prob = O.Problem(**p.kwargs)
    x = p.x0
    with CostProblem(**p.kwargs) as g:
        c, gr, t = g.cost_grad(x)
    c_es, g_es, t_es = O.calc_cost(x, prob, **ES)
    print("3C273 cost", c, c_es, abs(c - c_es) / c_es, "grad", rel_l2(gr, g_es))
    assert abs(c - c_es) <= 1e-11 * abs(c_es)
    assert rel_l2(gr, g_es) <= 1e-10
    for k in ("chisq_fcv", "chisq_cp", "chisq_ca", "l1", "tsv", "kl"):
        assert t[k] == pytest.approx(t_es[k], rel=1e-10)


def test_imaging_on_real_data_matches_oracle_solve(gpu_lib):
    from scipy.optimize import minimize

    from smili_amd.imaging import lbfgs
    from smili_amd.imaging.prep import prepare
    vis, bs, ca = load_tables()
    img = _setup()
    h = img.header
    out, info = lbfgs.imaging(img, vistable=vis, bstable=bs, catable=ca, niter=20, verbose=False,
                              return_info=True, **REGS)
    p = prepare(img.data[0, 0], np.deg2rad(h["dx"]), np.deg2rad(h["dy"]), h["nxref"], h["nyref"],
                vistable=vis, bstable=bs, catable=ca, **REGS)
    prob = O.Problem(**p.kwargs)

    def fg(xx):
        c, g, _ = O.calc_cost(xx, prob, **ES)
        return c, g

    res = minimize(fg, p.x0, jac=True, method="L-BFGS-B", bounds=[(0.0, None)] * p.x0.size,
                   options=dict(maxcor=lbfgs.lbfgsbprms["m"], ftol=lbfgs.lbfgsbprms["factr"] * np.finfo(float).eps,
                                gtol=0.0, maxiter=21, maxfun=10 ** 9, maxls=20))
    ref = (res.x / p.inormfactr).reshape(256, 256)
    got = np.asarray(out.data)[0, 0]
    err = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))
    print("3C273 imaging nit", info.niter, res.nit, "NRMSE", err, "cost", info.cost, res.fun)
    assert info.niter == res.nit
    assert err <= 1e-6
