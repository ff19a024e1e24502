"""Closed-form checks of the oracle's costs that finite differences cannot
see (they test a gradient against its cost, not the cost itself):

* second moment (image.f90:964-1138): an elliptical Gaussian image whose
  covariance is the prior (sm_maj, sm_min, sm_phi) has sm cost ~0, at two
  position angles, while the same image against a prior rotated by 90
  degrees has a cost of 2 (maj - min)^2 in closed form -- a swapped sine /
  cosine in Sigma_prior or a transposed moment fails;
* centroid (image.f90:793-863): an image symmetric about (Nxref, Nyref)
  has cost sqrt(1e-10) exactly and zero gradient; the same image shifted
  by (+2, -1) pixels has cost sqrt(5 + 1e-10) and a gradient that points
  back (a wrong offset sign fails);
* total flux (image.f90:773-787): cost (sum I - F)^2;
* the four reported chi-square terms (fftlib.f90:342-366): data exactly
  one standard deviation off the model give chisq_fcv = chisq_amp =
  chisq_cp = chisq_ca = 1 (a dropped factor 2 in the fcv normalisation,
  or the wrong count, fails), and the total is the weighted sum.
These run on the numpy oracle and, where it is built, on the reference's
own compiled image.f90 (oracle/_ref)."""
import math

import numpy as np
import pytest

from oracle import ref as R
from oracle import smili_oracle as O
from tests.helpers import random_problem

HAVE_REF = R.load() is not None


def _grid(Nx, Ny):
    yy, xx = np.mgrid[0:Ny, 0:Nx]
    return (xx.ravel() + 1).astype(np.int32), (yy.ravel() + 1).astype(np.int32)


def _gauss(Nx, Ny, xc, yc, S):
    """Unnormalised Gaussian with covariance S = [[Sxx, Sxy], [Sxy, Syy]]
    (pixels^2) centred at 1-based (xc, yc), x fastest."""
    xidx, yidx = _grid(Nx, Ny)
    d = np.stack([xidx - xc, yidx - yc])
    Si = np.linalg.inv(np.array([[S[0], S[2]], [S[2], S[1]]]))
    q = np.einsum("in,ij,jn->n", d, Si, d)
    return np.exp(-0.5 * q), xidx, yidx


def _sm_cost(x, xidx, yidx, Nx, Ny, maj, mn, phi):
    c_o, _ = O.calc_smreg(x, xidx, yidx, maj, mn, phi)
    out = [c_o]
    if HAVE_REF:
        c_r, _, t_r, _ = R.calc_cost_reg(x, xidx, yidx, Nx, Ny, sm_l=1.0, sm_maj=maj, sm_min=mn, sm_phi=phi)
        out.append(t_r["sm"])
    return out


@pytest.mark.parametrize("phi", [0.3, 1.2])
def test_second_moment_of_matching_gaussian_is_zero(phi):
    maj, mn = 9.0, 4.0                       # variances along the axes (pixels^2)
    Sp = O._sigma_prior(maj, mn, phi)
    # Sp is the covariance of the prior ellipse: its eigenvalues are maj, min
    ev = np.linalg.eigvalsh(np.array([[Sp[0], Sp[2]], [Sp[2], Sp[1]]]))
    assert np.allclose(sorted(ev), [mn, maj], rtol=1e-14)
    Nx = Ny = 64
    x, xidx, yidx = _gauss(Nx, Ny, 32.5, 30.0, Sp)
    for c in _sm_cost(x, xidx, yidx, Nx, Ny, maj, mn, phi):
        assert c <= 1e-20, c
    # the prior turned by 90 degrees: covariance error dS = (min - maj) R,
    # R = [[cos 2phi, sin 2phi], [sin 2phi, -cos 2phi]] -> cost = dSxx^2 +
    # dSyy^2 + 2 dSxy^2 = 2 (maj - min)^2
    for c in _sm_cost(x, xidx, yidx, Nx, Ny, maj, mn, phi + math.pi / 2):
        assert c == pytest.approx(2.0 * (maj - mn) ** 2, rel=1e-12)


def test_centroid_of_symmetric_image():
    rng = np.random.default_rng(3)
    Nx, Ny = 17, 13
    xidx, yidx = _grid(Nx, Ny)
    half = rng.uniform(0.1, 1.0, (Ny, Nx))
    img = half + half[::-1, ::-1]            # symmetric about the centre pixel (9, 7)
    x = img.ravel()
    for alpha in (1.0, 2.0):
        c, g = O.calc_cenreg(x, xidx, yidx, (Nx + 1) / 2, (Ny + 1) / 2, alpha)
        assert c == pytest.approx(math.sqrt(1e-10), rel=1e-9)
        # the same image about a reference pixel (2, -1) off its centre
        c2, g2 = O.calc_cenreg(x, xidx, yidx, (Nx + 1) / 2 - 2.0, (Ny + 1) / 2 + 1.0, alpha)
        assert c2 == pytest.approx(math.sqrt(5.0 + 1e-10), rel=1e-12)
        # moving flux towards the reference pixel (-x, +y) lowers the cost
        assert np.sum(g2 * (xidx - xidx.mean())) > 0 and np.sum(g2 * (yidx - yidx.mean())) < 0
        if HAVE_REF:
            cr, gr = R.calc_cenreg(x, xidx, yidx, (Nx + 1) / 2 - 2.0, (Ny + 1) / 2 + 1.0, alpha)
            assert cr == pytest.approx(c2, rel=1e-14)
            assert np.linalg.norm(gr - g2) <= 1e-13 * np.linalg.norm(gr)


def test_total_flux():
    x = np.linspace(0.1, 1.0, 20)
    c, g = O.calc_tfdreg(x, 7.0)
    assert c == pytest.approx((x.sum() - 7.0) ** 2, rel=1e-15) and g == pytest.approx(2 * (x.sum() - 7.0))


def one_sigma_problem(rng, fwd=O.dft_fwd):
    """Data one standard deviation off the model of x (fwd: the exact DFT,
    or the ES-kernel NUFFT for the GPU test): fcv off by (s, s) in (re, im),
    amplitudes by s, log closure amplitudes by s, closure phases by r with
    2 (1 - cos r) = s^2."""
    p, x = random_problem(rng, 20, 16, M=300, nfcv=120, namp=90, ncp=70, nca=60, noise=0.0)
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny)
    V = fwd(I2d, p.u, p.v)

    def model(idx):
        j = np.abs(idx) - 1
        return np.where(idx > 0, V[j], np.conj(V[j]))

    s = np.sqrt(p.Varfcv)
    p.Vfcv = model(p.uvidxfcv) + s * (1 + 1j) * np.where(rng.uniform(size=s.size) < 0.5, 1, -1)
    s = np.sqrt(p.Varamp)
    p.Vamp = np.abs(model(p.uvidxamp)) + s * np.where(rng.uniform(size=s.size) < 0.5, 1, -1)
    s = np.sqrt(p.Varca)
    a = [np.log(np.abs(model(p.uvidxca[k]))) for k in range(4)]
    p.CA = a[0] + a[1] - a[2] - a[3] + s * np.where(rng.uniform(size=s.size) < 0.5, 1, -1)
    s = np.sqrt(p.Varcp)
    r = np.arccos(1.0 - s ** 2 / 2.0)
    ph = sum(np.angle(model(p.uvidxcp[k])) for k in range(3))
    p.CP = ph + r * np.where(rng.uniform(size=s.size) < 0.5, 1, -1)
    return p, x


def test_chisq_terms_are_one_at_one_sigma():
    rng = np.random.default_rng(17)
    p, x = one_sigma_problem(rng)
    c, g, t = O.calc_cost(x, p)
    for k in ("chisq_fcv", "chisq_amp", "chisq_cp", "chisq_ca"):
        assert t[k] == pytest.approx(1.0, abs=1e-9), k
    n = dict(fcv=len(p.uvidxfcv), amp=len(p.uvidxamp), cp=p.uvidxcp.shape[1], ca=p.uvidxca.shape[1])
    # the total chi-square is the weighted sum with the data counts of calc_chisq
    tot = 2 * n["fcv"] * p.wfcv + n["amp"] * p.wamp + n["cp"] * p.wcp + n["ca"] * p.wca
    assert t["chisq"] == pytest.approx(tot, rel=1e-9)
