"""Pulse-function support of imaging(pulse_fwhm=...) (lbfgs.py:436-478,
574-575, 1114-1194), against closed forms:

* the pulse's Fourier transform K(rho) = exp(-pi^2 theta^2 rho^2 / (4 ln 2))
  equals the direct DFT of a finely sampled unit-flux Gaussian of FWHM theta;
* a Gaussian sky of FWHM a seen through the pulse has visibilities of a
  Gaussian of FWHM sqrt(a^2 + theta^2); deburring its table gives the sky's
  own visibilities, amplitudes and sigmas divided by K, closure amplitudes
  shifted by log(K1 K2 / (K3 K4)), closure phases' sigma / amp unchanged;
* the output convolution keeps the total flux, adds the kernel's discrete
  variance to the image's and moves its centroid by the kernel's mean
  offset from its origin (element n // 2)."""
import math

import numpy as np
import pandas as pd
import pytest

from smili_amd.imaging import deburr as D

UAS = math.radians(1.0 / 3.6e9)


def _gauss_vis(fwhm_rad, u, v, flux=1.0):
    return flux * D.pulse_ft(fwhm_rad, np.hypot(u, v))


def test_pulse_ft_is_the_gaussian_dft():
    n, dx = 128, 1.0 * UAS
    theta = 12 * UAS
    x = (np.arange(n) - n / 2) * dx
    X, Y = np.meshgrid(x, x)
    s = theta / (2 * math.sqrt(2 * math.log(2)))
    img = np.exp(-(X ** 2 + Y ** 2) / (2 * s * s))
    img /= img.sum()
    rng = np.random.default_rng(0)
    rho = rng.uniform(0, 2e10, 50)
    ang = rng.uniform(0, 2 * np.pi, 50)
    u, v = rho * np.cos(ang), rho * np.sin(ang)
    # direct sum of the image's visibilities at (u, v) in wavelengths
    V = np.array([np.sum(img * np.exp(2j * np.pi * (uu * X + vv * Y))) for uu, vv in zip(u, v)])
    assert np.max(np.abs(V.real - D.pulse_ft(theta, rho))) <= 1e-9
    assert np.max(np.abs(V.imag)) <= 1e-9


@pytest.mark.parametrize("as_frame", [False, True])
def test_deburred_tables_are_the_sky_tables(as_frame):
    rng = np.random.default_rng(1)
    a, theta = 20 * UAS, 8 * UAS
    both = math.hypot(a, theta)                 # sky (x) pulse: FWHMs add in quadrature
    m = 40
    u, v = rng.uniform(-8e9, 8e9, (2, m))
    obs_amp = _gauss_vis(both, u, v, 0.7)
    sig = rng.uniform(0.01, 0.02, m)
    vis = dict(u=u, v=v, uvdist=np.hypot(u, v), amp=obs_amp, phase=np.zeros(m), sigma=sig)
    vis_nouv = dict(u=u, v=v, amp=obs_amp, phase=np.zeros(m), sigma=sig)       # uvdist from u, v
    k = D.pulse_ft(theta, np.hypot(u, v))
    for t in (vis, vis_nouv):
        t = pd.DataFrame(t) if as_frame else t
        d = D.deburr_vistable(t, theta)
        np.testing.assert_allclose(np.asarray(d["amp"]), _gauss_vis(a, u, v, 0.7), rtol=1e-12)
        np.testing.assert_allclose(np.asarray(d["sigma"]), sig / k, rtol=1e-14)
        assert np.array_equal(np.asarray(t["amp"]), obs_amp)               # the input is untouched
    # closure tables over random triangles / quadrangles of baselines
    uv = rng.uniform(-8e9, 8e9, (8, m))
    bs = {"u12": uv[0], "v12": uv[1], "u23": uv[2], "v23": uv[3], "u31": uv[4], "v31": uv[5],
          "amp": rng.uniform(0.1, 1, m), "sigma": rng.uniform(0.01, 0.1, m), "phase": rng.uniform(-90, 90, m)}
    for key in ("12", "23", "31"):
        bs["uvdist" + key] = np.hypot(bs["u" + key], bs["v" + key])
    bs = pd.DataFrame(bs) if as_frame else bs
    d = D.deburr_bstable(bs, theta)
    np.testing.assert_allclose(np.asarray(d["sigma"]) / np.asarray(d["amp"]),
                               np.asarray(bs["sigma"]) / np.asarray(bs["amp"]), rtol=1e-14)
    assert np.array_equal(np.asarray(d["phase"]), np.asarray(bs["phase"]))
    ca = {f"u{n}": uv[2 * n - 2] for n in range(1, 5)}
    ca.update({f"v{n}": uv[2 * n - 1] for n in range(1, 5)})
    # closure amplitudes of the observed (sky (x) pulse) visibilities
    amps = [_gauss_vis(both, ca[f"u{n}"], ca[f"v{n}"]) for n in range(1, 5)]
    ca.update(amp=amps[0] * amps[1] / amps[2] / amps[3], sigma=np.full(m, 0.05), logsigma=np.full(m, 0.05))
    ca["logamp"] = np.log(ca["amp"])
    ca = pd.DataFrame(ca) if as_frame else ca
    d = D.deburr_catable(ca, theta)
    sky = [_gauss_vis(a, uv[2 * n - 2], uv[2 * n - 1]) for n in range(1, 5)]
    np.testing.assert_allclose(np.asarray(d["logamp"]), np.log(sky[0] * sky[1] / sky[2] / sky[3]), atol=1e-12)
    np.testing.assert_allclose(np.asarray(d["amp"]), sky[0] * sky[1] / sky[2] / sky[3], rtol=1e-12)
    assert np.array_equal(np.asarray(d["logsigma"]), np.asarray(ca["logsigma"]))


@pytest.mark.parametrize("n", [32, 33])
def test_output_convolution_moments(n):
    rng = np.random.default_rng(n)
    img = np.zeros((n, n))
    img[10:20, 12:22] = rng.uniform(0, 1, (10, 10))
    fw = 3.0
    out = D.convolve_gauss(img, fw, save_totalflux=True)
    assert out.sum() == pytest.approx(img.sum(), rel=1e-13)
    # the sampled kernel and its moments about its origin (element n // 2)
    off = np.arange(n) - n // 2
    s = fw / math.sqrt(2 * math.log(2)) / 2
    k1 = np.exp(-(np.arange(n) - (n - 1) / 2.0) ** 2 / (2 * s * s))
    k1 /= k1.sum()
    km = np.sum(k1 * off)
    kv = np.sum(k1 * off ** 2) - km ** 2

    def moments(a, axis):
        p = a.sum(axis=axis)
        c = np.arange(n)
        m = np.sum(p * c) / p.sum()
        return m, np.sum(p * (c - m) ** 2) / p.sum()

    for axis in (0, 1):
        m0, v0 = moments(img, axis)
        m1, v1 = moments(out, axis)
        assert m1 == pytest.approx(m0 + km, abs=1e-9)
        assert v1 == pytest.approx(v0 + kv, rel=1e-9)
    # a point source becomes the kernel, centred on the source
    pt = np.zeros((n, n))
    pt[15, 16] = 2.0
    o = D.convolve_gauss(pt, fw, save_totalflux=True)
    k2 = np.outer(k1, k1) * 2.0
    ys, xs = np.arange(n), np.arange(n)
    expect = np.zeros((n, n))
    for y in ys:
        for x in xs:
            ky, kx = y - 15 + n // 2, x - 16 + n // 2
            if 0 <= ky < n and 0 <= kx < n:
                expect[y, x] = k2[ky, kx]
    expect *= 2.0 / expect.sum()
    np.testing.assert_allclose(o, expect, atol=1e-15)
