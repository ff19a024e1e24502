"""GPU transform and model routines (smili_amd.fftlib) against the FP64
ES-kernel oracle and the exact DFT: the f2py-convention wrappers of
src/fortlib/fftlib.f90:70-244 and model_fcv/amp/ca/cp (fftlib.f90:744-1254,
fftlib3d.f90:207-940).  FP64 NUFFT: transforms rel-l2 <= 1e-12 against the
ES oracle (same kernel) and <= 2e-6 against the DFT (the kernel's eps);
chi-square rel <= 1e-11, gradients rel-l2 <= 1e-10, model / residual
values rel-l2 <= 1e-11."""
import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import random_movie, random_problem, rel_l2

pytestmark = pytest.mark.gpu


def _raw_uv(rng, n, umax=2.5):
    # table coordinates of either sign (no uv-list flip)
    return rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)


@pytest.mark.parametrize("shape", [(16, 16), (24, 20), (33, 17)])
def test_nufft_wrappers(gpu_lib, shape):
    from smili_amd import fftlib as F
    rng = np.random.default_rng(shape[0])
    Nx, Ny = shape
    u, v = _raw_uv(rng, 300)
    I_xy = rng.uniform(0, 1, (Nx, Ny))                  # I2d(Nx,Ny) as f2py passes it
    V = F.nufft_fwd(u, v, I_xy)
    assert rel_l2(V, O.es_nufft_type2(I_xy.T, u, v)) <= 1e-12
    assert rel_l2(V, O.dft_fwd(I_xy.T, u, v)) <= 2e-6
    Vr, Vi = F.nufft_fwd_real(u, v, I_xy)
    assert np.array_equal(Vr + 1j * Vi, V)
    a, b = rng.standard_normal(300), rng.standard_normal(300)
    A = O.es_nufft_type1(a + 1j * b, u, v, Nx, Ny)      # [Ny, Nx]
    assert rel_l2(F.nufft_adj_real1d(u, v, a, b, Nx, Ny), A.real.reshape(-1)) <= 1e-12
    re, im = F.nufft_adj_real(u, v, a, b, Nx, Ny)
    assert rel_l2(re, A.real.reshape(-1)) <= 1e-12 and rel_l2(im, A.imag.reshape(-1)) <= 1e-12
    C = F.nufft_adj(u, v, a + 1j * b, Nx, Ny)
    assert C.shape == (Nx, Ny) and rel_l2(C, A.T) <= 1e-12
    assert rel_l2(C, O.dft_adj(a + 1j * b, u, v, Nx, Ny).T) <= 2e-6
    c2 = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    R = F.nufft_adj_resid(u, v, a + 1j * b, c2, Nx, Ny)
    Rref = O.nufft_adj_resid(a + 1j * b, c2, u, v, Nx, Ny, O.es_nufft_type1)
    assert rel_l2(R, Rref.T) <= 1e-12
    assert rel_l2(F.nufft_adjrea(u, v, c2, Nx, Ny), O.es_nufft_type1(c2, u, v, Nx, Ny).real.T) <= 1e-12


def _oracle_model(p, x, kind, **blk):
    """chi-square (weight 1), gradient, model and residual of one data block
    from the FP64 ES-kernel oracle, following fftlib.f90:744-1254."""
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny)
    V = O.es_nufft_type2(I2d, p.u, p.v)
    M = p.u.size
    r, i = np.zeros(M), np.zeros(M)
    if kind == "fcv":
        c = O.chisq_fcv(V, blk["idx"], blk["obs"], blk["var"], 1.0, r, i)
    elif kind == "amp":
        c = O.chisq_amp(V, blk["idx"], blk["obs"], blk["var"], 1.0, r, i)
    elif kind == "cp":
        c = O.chisq_cp(V, blk["idx"], blk["obs"], blk["var"], 1.0, r, i)
    else:
        c = O.chisq_ca(V, blk["idx"], blk["obs"], blk["var"], 1.0, r, i)
    g = O.nufft_adj_resid(r, i, p.u, p.v, p.Nx, p.Ny, O.es_nufft_type1)[p.yidx - 1, p.xidx - 1]
    return c, g, V


@pytest.mark.parametrize("window", [False, True])
def test_model_functions(gpu_lib, window):
    from smili_amd import fftlib as F
    rng = np.random.default_rng(5 + window)
    p, x = random_problem(rng, 20, 18, M=250, window=window)
    p.Nxref, p.Nyref = 7.0, 12.5                       # off-centre reference pixel
    geom = (x, p.xidx, p.yidx, p.Nxref, p.Nyref, p.Nx, p.Ny, p.u, p.v)
    # fcv: data in the reference-pixel phase; model / resid returned likewise
    idx = p.uvidxfcv
    us, vs = p.u[np.abs(idx) - 1] * np.sign(idx), p.v[np.abs(idx) - 1] * np.sign(idx)
    raw = F.phashift_c2r(us, vs, p.Nxref, p.Nyref, p.Nx, p.Ny, p.Vfcv)
    chisq, grad, mr, mi, rr, ri = F.model_fcv(*geom, idx, raw.real, raw.imag, p.Varfcv)
    c, g, V = _oracle_model(p, x, "fcv", idx=idx, obs=p.Vfcv, var=p.Varfcv)
    model = np.where(idx > 0, V[np.abs(idx) - 1], np.conj(V[np.abs(idx) - 1]))
    model_raw = F.phashift_c2r(us, vs, p.Nxref, p.Nyref, p.Nx, p.Ny, model)
    assert abs(chisq - c) <= 1e-11 * c and rel_l2(grad, g) <= 1e-10
    assert rel_l2(mr + 1j * mi, model_raw) <= 1e-11
    assert rel_l2(rr + 1j * ri, raw - model_raw) <= 1e-9
    # amp
    chisq, grad, m, r = F.model_amp(*geom, p.uvidxamp, p.Vamp, p.Varamp)
    c, g, V = _oracle_model(p, x, "amp", idx=p.uvidxamp, obs=p.Vamp, var=p.Varamp)
    assert abs(chisq - c) <= 1e-11 * c and rel_l2(grad, g) <= 1e-10
    assert rel_l2(m, np.abs(V[np.abs(p.uvidxamp) - 1])) <= 1e-11
    assert rel_l2(r, p.Vamp - m) == 0.0
    # closure phase: model and residual wrapped to (-pi, pi]
    chisq, grad, m, r = F.model_cp(*geom, p.uvidxcp, p.CP, p.Varcp)
    c, g, V = _oracle_model(p, x, "cp", idx=p.uvidxcp, obs=p.CP, var=p.Varcp)
    ph = sum(np.angle(V[np.abs(p.uvidxcp[k]) - 1]) * np.sign(p.uvidxcp[k]) for k in range(3))
    assert abs(chisq - c) <= 1e-11 * c and rel_l2(grad, g) <= 1e-10
    assert np.all(np.abs(m) <= np.pi) and np.all(np.abs(r) <= np.pi)
    assert np.abs(np.angle(np.exp(1j * (m - ph)))).max() <= 1e-9
    assert np.abs(np.angle(np.exp(1j * (r - (p.CP - ph))))).max() <= 1e-9
    # log closure amplitude
    chisq, grad, m, r = F.model_ca(*geom, p.uvidxca, p.CA, p.Varca)
    c, g, V = _oracle_model(p, x, "ca", idx=p.uvidxca, obs=p.CA, var=p.Varca)
    la = [np.log(np.abs(V[np.abs(p.uvidxca[k]) - 1])) for k in range(4)]
    assert abs(chisq - c) <= 1e-11 * c and rel_l2(grad, g) <= 1e-10
    assert rel_l2(m, la[0] + la[1] - la[2] - la[3]) <= 1e-11


def test_model_fcv_movie(gpu_lib):
    """fftlib3d.model_fcv: per-frame transforms of the concatenated uv list."""
    from smili_amd import fftlib as F
    rng = np.random.default_rng(9)
    mv, x = random_movie(rng, 3, 16, 16, [80, 0, 60], namp=0, ncp=0, nca=0)
    p = mv.frame
    idx = p.uvidxfcv
    us, vs = p.u[np.abs(idx) - 1] * np.sign(idx), p.v[np.abs(idx) - 1] * np.sign(idx)
    raw = F.phashift_c2r(us, vs, p.Nxref, p.Nyref, p.Nx, p.Ny, p.Vfcv)
    chisq, grad, mr, mi, rr, ri = F.fftlib3d.model_fcv(x, p.xidx, p.yidx, p.Nxref, p.Nyref, p.Nx, p.Ny, 3,
                                                       p.u, p.v, mv.Nuvs, idx, raw.real, raw.imag, p.Varfcv)
    # oracle: calc_chisq3d with weight 1 and only the fcv block
    import dataclasses
    f1 = dataclasses.replace(p, wfcv=1.0)
    m1 = O.Movie(frame=f1, Nuvs=mv.Nuvs)
    c, g, t, V = O.calc_chisq3d(x, m1, fwd=O.es_nufft_type2, adj=O.es_nufft_type1)
    assert abs(chisq - c) <= 1e-11 * c and rel_l2(grad, g) <= 1e-10
    model = np.where(idx > 0, V[np.abs(idx) - 1], np.conj(V[np.abs(idx) - 1]))
    assert rel_l2(mr + 1j * mi, F.phashift_c2r(us, vs, p.Nxref, p.Nyref, p.Nx, p.Ny, model)) <= 1e-11
