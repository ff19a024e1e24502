"""Visibility-table operations on the GPU transforms (smili_amd.uvtable)
against direct FP64 sums: model / residual tables (vistable.py:361-433),
dirty beam and residual map (vistable.py:2627-2778).  ES-kernel eps 1e-6:
rel-l2 <= 2e-6 against the direct sums."""
import numpy as np
import pytest

from smili_amd import synth
from smili_amd.imaging.image import Image, Movie
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
UAS = synth.UAS


def _setup(seed=0, n=32, nvis=500):
    rng = np.random.default_rng(seed)
    img = Image(n, n, 2.0, nxref=n / 2 + 3.0, nyref=n / 2 - 2.0)
    sky = synth.random_gaussians(rng, 3, 5, total_flux=1.0, fov=40 * UAS)
    dx, dy = np.deg2rad(img.header["dx"]), np.deg2rad(img.header["dy"])
    img.data[0, 0] = synth.sky_image(sky, n, n, dx, dy, img.header["nxref"], img.header["nyref"])
    u = rng.uniform(-8e9, 8e9, nvis)
    v = rng.uniform(-8e9, 8e9, nvis)
    tab = synth.vis_table(rng, sky, u, v, 0.01)
    return img, tab


def _pixel_positions(img):
    h = img.header
    x = np.deg2rad(h["dx"]) * (np.arange(h["nx"]) + 1 - h["nxref"])
    y = np.deg2rad(h["dy"]) * (np.arange(h["ny"]) + 1 - h["nyref"])
    return np.meshgrid(x, y)


def _dft_model(img, u, v):
    X, Y = _pixel_positions(img)
    I = img.data[0, 0]
    ph = 2 * np.pi * (np.outer(u, X.reshape(-1)) + np.outer(v, Y.reshape(-1)))
    return np.exp(1j * ph) @ I.reshape(-1)


def test_eval_and_residual_tables(gpu_lib):
    from smili_amd import uvtable as T
    img, tab = _setup()
    Vm = _dft_model(img, tab["u"], tab["v"])
    mt = T.eval_image(tab, img)
    assert rel_l2(T.comp(mt), Vm) <= 2e-6
    rt = T.residual_image(tab, img)
    assert rel_l2(T.comp(rt), T.comp(tab) - Vm) <= 2e-4        # residual ~1% of the data
    chisq, rchisq = T.chisq_image(tab, img)
    ref = np.sum(np.abs(T.comp(tab) - Vm) ** 2 / tab["sigma"] ** 2)
    assert chisq == pytest.approx(ref, rel=1e-4) and rchisq == pytest.approx(chisq / (2 * len(tab["u"])))
    at = T.eval_image(tab, img, amptable=True)
    assert rel_l2(at["amp"], np.abs(Vm)) <= 2e-6 and np.all(at["phase"] == 0)


def test_movie_tables_follow_rows(gpu_lib):
    """A movie's model table maps every row to its own frame, whatever the
    table's row order."""
    from smili_amd import uvtable as T
    img, tab = _setup(1)
    rng = np.random.default_rng(2)
    Nt = 3
    frames = [img.copy() for _ in range(Nt)]
    for k, f in enumerate(frames):
        f.data[0, 0] *= 1 + 0.3 * k
    tab["frmidx"] = rng.integers(0, Nt, len(tab["u"]))
    mov = Movie(frames)
    mt = T.eval_image(tab, mov)
    Vm = _dft_model(img, tab["u"], tab["v"]) * (1 + 0.3 * tab["frmidx"])
    assert rel_l2(T.comp(mt), Vm) <= 2e-6


def test_dirty_beam_and_residual_map(gpu_lib):
    from smili_amd import uvtable as T
    img, tab = _setup(3, n=24, nvis=300)
    X, Y = _pixel_positions(img)
    w = T.uvweight(tab, uniform=False, werror=0)
    beam = T.map_beam(tab, img, uniform=False, normalize=False)
    ph = 2 * np.pi * (np.outer(tab["u"], X.reshape(-1)) + np.outer(tab["v"], Y.reshape(-1)))
    ref = (w @ np.cos(ph)).reshape(X.shape)                   # sum over (u, -u) of w/2 cos
    assert rel_l2(beam.data[0, 0], ref) <= 2e-6
    nb = T.map_beam(tab, img, uniform=False)
    assert nb.data[0, 0].max() == pytest.approx(1.0)
    iy, ix = np.unravel_index(np.argmax(nb.data[0, 0]), X.shape)
    assert (ix + 1, iy + 1) == (img.header["nxref"], img.header["nyref"])
    res = T.map_residual(tab, img, werror=-2)
    R = T.comp(T.residual_image(tab, img))
    wr = T.uvweight(tab, uniform=False, werror=-2)
    ref = (wr @ (R[:, None] * np.exp(-1j * ph)).real).reshape(X.shape)
    assert rel_l2(res.data[0, 0], ref) <= 2e-5
    cl = T.map_clean(tab, img)
    assert np.allclose(cl.data[0, 0], res.data[0, 0] + img.data[0, 0])
