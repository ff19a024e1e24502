"""Host parts of the visibility-table operations (smili_amd.uvtable) on
CPU: uv density against the reference's brute-force count
(smili/uvdata/uvtable/vistable.py:2432-2477), weight normalisation, and the
beam fit against closed forms."""
import numpy as np
import pytest

from smili_amd import uvtable as T
from smili_amd.imaging.image import Image


def _brute_density(u, v, npix):
    out = []
    for i in range(u.size):
        a = ((u - u[i]) ** 2 + (v - v[i]) ** 2) < npix ** 2
        b = ((u + u[i]) ** 2 + (v + v[i]) ** 2) < npix ** 2
        out.append(np.count_nonzero(a | b))
    return np.asarray(out, float)


def test_uvdensity_matches_brute_force_count():
    rng = np.random.default_rng(0)
    img = Image(32, 32, 2.0)
    n = 400
    u = rng.normal(0, 3e9, n)
    v = rng.normal(0, 3e9, n)
    u[:20] = -u[20:40]           # conjugate pairs
    v[:20] = -v[20:40]
    tab = dict(u=u, v=v, amp=np.ones(n), phase=np.zeros(n), sigma=np.ones(n))
    x, y = T._xygrid_rad(img)
    du = 1 / (np.sqrt(x * x + y * y).max() * 2)
    for npix in (2, 5):
        np.testing.assert_array_equal(T.uvdensity(tab, img, npix), _brute_density(u / du, v / du, npix))


def test_uvweight_normalised_and_error_power():
    rng = np.random.default_rng(1)
    n = 100
    sig = rng.uniform(0.5, 2, n)
    tab = dict(u=rng.normal(0, 1e9, n), v=rng.normal(0, 1e9, n), amp=np.ones(n), phase=np.zeros(n), sigma=sig)
    w = T.uvweight(tab, uniform=False, werror=-2)
    assert w.sum() == pytest.approx(1.0)
    np.testing.assert_allclose(w, sig ** -2 / np.sum(sig ** -2))
    wu = T.uvweight(tab, uniform=True, image=Image(16, 16, 2.0), werror=0)
    assert wu.sum() == pytest.approx(1.0) and np.all(wu > 0)


def test_fit_beam_closed_forms():
    """Points on an ellipse |u| = a, |v| = b (axis-aligned): second moments
    a^2/2, b^2/2, no cross term; FWHMs fudge / sqrt(var), pa 0 or -90."""
    n = 720
    t = np.linspace(0, 2 * np.pi, n, endpoint=False)
    a, b = 4e9, 1e9
    tab = dict(u=a * np.cos(t), v=b * np.sin(t), amp=np.ones(n), phase=np.zeros(n), sigma=np.ones(n))
    bm = T.fit_beam(tab, Image(16, 16, 2.0), uniform=False, angunit="uas")
    fudge = np.sqrt(8 * np.log(2)) / (2 * np.pi)
    rad2uas = 180 / np.pi * 3600e6
    assert bm["minsize"] == pytest.approx(fudge / np.sqrt(a * a / 2) * rad2uas, rel=1e-9)
    assert bm["majsize"] == pytest.approx(fudge / np.sqrt(b * b / 2) * rad2uas, rel=1e-9)
    assert abs(bm["pa"]) < 1e-6
