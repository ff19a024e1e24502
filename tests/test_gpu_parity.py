"""GPU parity of the HIP cost against the FP64 oracle, through the C ABI.

The GPU runs the ES-kernel NUFFT (width 7, eps 1e-6) with FP32 spreading,
interpolation and FFT, and FP64 chi-square / regularizer arithmetic.  Two
oracles check it:

* the FP64 ES-kernel NUFFT oracle (same approximation, FP64 arithmetic):
  isolates the FP32 error of the GPU path --
  cost rel <= 1e-6, model visibilities rel-l2 <= 1e-6, gradient rel-l2 <=
  5e-6 + 0.5 e_es (the closure terms' 1/|V|^2 factors amplify the FP32
  rounding of V exactly as they amplify the kernel error e_es below);
* the exact FP64 direct DFT (what the ES kernel approximates): the GPU must
  be as close to it as the FP64 ES oracle is, up to the FP32 error --
  visibilities / adjoint rel-l2 <= 4e-6, cost rel <= 2e-6 + 2 e_es,
  gradient rel-l2 <= 5e-6 + 1.5 e_es, where e_es is the FP64 ES oracle's own
  error against the DFT on the same case (1e-6 .. 2e-5 here: the 1/|V|^2
  factors of the closure terms amplify the kernel's eps).
Regularizer terms are FP64 end to end: rel <= 1e-12.
"""
import math

import numpy as np
import pytest

from oracle import smili_oracle as O
from tests.helpers import problem_kwargs, random_problem, rel_l2

pytestmark = pytest.mark.gpu

TOL_V = 4e-6
TOL_V_ES = 1e-6
TOL_COST_ES = 1e-6
TOL_GRAD_ES = 5e-6
TOL_REG = 1e-12


def _mk(p, tune=None):
    """FP32 NUFFT unless the test asks otherwise (these tolerances are the
    FP32 path's; the FP64 path is checked at rounding level below)."""
    from smili_amd.cost import CostProblem
    t = {"precision": 32}
    t.update(tune or {})
    return CostProblem(**problem_kwargs(p), tune=t)


SHAPES = [(16, 16), (24, 20), (17, 13), (32, 48), (64, 64), (40, 30)]


@pytest.mark.parametrize("shape", SHAPES)
def test_type2_matches_dft(gpu_lib, shape):
    rng = np.random.default_rng(100 + shape[0])
    p, x = random_problem(rng, shape[0], shape[1], M=400)
    with _mk(p) as g:
        V = g.model_vis(x)
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny)
    Vd = O.dft_fwd(I2d, p.u, p.v)
    Ve = O.es_nufft_type2(I2d, p.u, p.v)
    err = rel_l2(V, Vd)
    print("type2", shape, "vs DFT", err, "vs ES64", rel_l2(V, Ve))
    assert err <= TOL_V
    assert rel_l2(V, Ve) <= TOL_V_ES


@pytest.mark.parametrize("shape", SHAPES)
def test_type1_matches_dft(gpu_lib, shape):
    rng = np.random.default_rng(200 + shape[0])
    p, x = random_problem(rng, shape[0], shape[1], M=400)
    a = rng.standard_normal(p.u.size)
    b = rng.standard_normal(p.u.size)
    with _mk(p) as g:
        out = g.adjoint_resid(a, b)
    ref = O.nufft_adj_resid(a, b, p.u, p.v, p.Nx, p.Ny)[p.yidx - 1, p.xidx - 1]
    err = rel_l2(out, ref)
    print("type1", shape, err)
    assert err <= TOL_V


def _check_cost(p, x, g):
    c_ref, g_ref, t_ref = O.calc_cost(x, p)
    c_es, g_es, t_es = O.calc_cost(x, p, fwd=O.es_nufft_type2, adj=O.es_nufft_type1)
    c, gr, t = g.cost_grad(x)
    e_es_c = abs(c_es - c_ref) / abs(c_ref)
    e_es_g = rel_l2(g_es, g_ref)
    print("vs ES64: cost %.2e grad %.2e | vs DFT: cost %.2e grad %.2e | ES64 vs DFT: cost %.2e grad %.2e" % (
        abs(c - c_es) / abs(c_es), rel_l2(gr, g_es), abs(c - c_ref) / abs(c_ref), rel_l2(gr, g_ref),
        e_es_c, e_es_g))
    assert abs(c - c_es) <= TOL_COST_ES * abs(c_es)
    assert rel_l2(gr, g_es) <= TOL_GRAD_ES + 0.5 * e_es_g
    assert abs(c - c_ref) <= (2e-6 + 2 * e_es_c) * abs(c_ref)
    assert rel_l2(gr, g_ref) <= 5e-6 + 1.5 * e_es_g
    for k in ("chisq_fcv", "chisq_amp", "chisq_cp", "chisq_ca"):
        if t_es[k] != 0:
            assert abs(t[k] - t_es[k]) <= 1e-5 * abs(t_es[k]), k
    for k in ("l1", "tv", "tsv", "kl", "gs", "tfd", "cen", "sm"):
        if t_ref[k] != 0:
            assert abs(t[k] - t_ref[k]) <= TOL_REG * abs(t_ref[k]) + 1e-300, k
        else:
            assert t[k] == 0.0, k


@pytest.mark.parametrize("shape,window", [((24, 20), False), ((32, 32), True), ((17, 13), False),
                                          ((64, 48), True)])
def test_cost_grad_all_terms(gpu_lib, shape, window):
    rng = np.random.default_rng(7 + shape[0] + window)
    p, x = random_problem(rng, shape[0], shape[1], M=500, window=window)
    with _mk(p) as g:
        _check_cost(p, x, g)


@pytest.mark.parametrize("only", ["fcv", "amp", "cp", "ca"])
def test_single_data_type(gpu_lib, only):
    rng = np.random.default_rng({"fcv": 1, "amp": 2, "cp": 3, "ca": 4}[only])
    n = dict(nfcv=0, namp=0, ncp=0, nca=0)
    n["n" + only] = 300
    p, x = random_problem(rng, 20, 20, M=250, regs=(), **n)
    with _mk(p) as g:
        _check_cost(p, x, g)


def test_log_regularizers_across_decades(gpu_lib):
    """MEM and GS take log(|I|_eps / prior) pixel by pixel; with priors over
    500 decades and pixels over 6, and the regularizers weighted to dominate
    the gradient, the kernels' logarithm (a series of its own, not the
    library's) must give the oracle's (numpy's log) gradient and terms to
    rounding (image.f90:639-709)."""
    import dataclasses
    rng = np.random.default_rng(21)
    p, x = random_problem(rng, 20, 20, M=150, regs=("kl", "gs"))
    n = x.size
    p = dataclasses.replace(p, kl_l=1e3, gs_l=1e3, kl_wgt=10.0 ** rng.uniform(-250, 250, n),
                            gs_wgt=10.0 ** rng.uniform(-250, 250, n))
    x = x * 10.0 ** rng.uniform(-3, 3, n)
    with _mk(p, tune={"precision": 64}) as g:
        c, gr, t = g.cost_grad(x)
    c_es, g_es, t_es = O.calc_cost(x, p, fwd=O.es_nufft_type2, adj=O.es_nufft_type1)
    print("grad rel-l2 %.2e, max per pixel %.2e" % (rel_l2(gr, g_es),
                                                    np.max(np.abs(gr - g_es) / np.abs(g_es))))
    assert rel_l2(gr, g_es) <= 1e-13
    assert np.all(np.abs(gr - g_es) <= 1e-12 * np.abs(g_es))
    for k in ("kl", "gs"):
        assert abs(t[k] - t_es[k]) <= 1e-13 * abs(t_es[k]), k


def test_moment_regularizers(gpu_lib):
    rng = np.random.default_rng(11)
    p, x = random_problem(rng, 24, 24, M=200, extra_regs=("tfd", "cen", "sm"))
    with _mk(p) as g:
        _check_cost(p, x, g)


def test_tv_reset_quirk_on_flat_patch(gpu_lib):
    """A flat patch makes neighbour TV terms vanish: the reference resets
    the accumulated gradient to 0 there (image.f90:577-578, 603-604)."""
    rng = np.random.default_rng(5)
    p, x = random_problem(rng, 16, 16, M=100, regs=("tv",))
    x2 = O.image_1d_to_2d(x, p.xidx, p.yidx, 16, 16)
    x2[4:9, 4:9] = 1.0
    x = x2[p.yidx - 1, p.xidx - 1]
    with _mk(p) as g:
        _check_cost(p, x, g)


def test_graph_and_eager_agree_and_repeat(gpu_lib):
    rng = np.random.default_rng(9)
    p, x = random_problem(rng, 32, 32, M=600)
    with _mk(p, tune=dict(use_graph=1)) as g1, _mk(p) as g2:
        c1, gr1, _ = g1.cost_grad(x)
        c1b, gr1b, _ = g1.cost_grad(x)
        c2, gr2, _ = g2.cost_grad(x)
        assert g1.info()["graph"] == 1 and g2.info()["graph"] == 0
    # the cost is a fixed-order reduction: bitwise reproducible
    assert c1 == c1b == c2
    # the gradient goes through float atomics: reproducible to rounding
    assert rel_l2(gr1b, gr1) < 1e-6 and rel_l2(gr2, gr1) < 1e-6


@pytest.mark.parametrize("shape,window,prec", [((24, 20), False, 32), ((32, 32), True, 64), ((64, 48), True, 32),
                                               ((17, 13), False, 64)])
def test_deterministic_mode_matches_oracle(gpu_lib, shape, window, prec):
    """smili_tune.deterministic: the same parity as the default path
    (all four data types, five regularizers, windows), in FP32 and FP64."""
    rng = np.random.default_rng(70 + shape[0] + window)
    p, x = random_problem(rng, shape[0], shape[1], M=500, window=window)
    with _mk(p, tune=dict(deterministic=1, precision=prec)) as g:
        _check_cost(p, x, g)


@pytest.mark.parametrize("precision", [32, 64])
def test_term_errors_against_dft(gpu_lib, precision):
    """Every output term against the exact-DFT oracle: the chi-square terms
    carry the ES kernel's error (eps 1e-6; FP32 grids add their rounding),
    the regularizers none (they do not touch the transforms)."""
    rng = np.random.default_rng(41 + precision)
    p, x = random_problem(rng, 64, 64, M=1500, nfcv=600, namp=500, ncp=400, nca=400, noise=0.02)
    with _mk(p, tune=dict(precision=precision)) as g:
        c, gr, t = g.cost_grad(x)
    c_ref, g_ref, t_ref = O.calc_cost(x, p)
    errs = {k: abs(t[k] - t_ref[k]) / abs(t_ref[k]) for k in t_ref if t_ref[k] != 0}
    print("precision", precision, " ".join("%s %.1e" % kv for kv in sorted(errs.items())))
    for k, e in errs.items():
        if k.startswith("chisq") or k == "cost":
            assert e <= (2e-5 if precision == 32 else 5e-6), (k, e)
        else:
            assert e <= 1e-12, (k, e)


@pytest.mark.parametrize("precision", [32, 64])
def test_distinct_point_chunks_match_oracle(gpu_lib, precision):
    """Visibilities and amplitudes that each sit at their own uv point take
    the chi-square kernel's direct path (no LDS staging of the points);
    both paths (the closures here repeat points) match the oracle."""
    rng = np.random.default_rng(31)
    p, x = random_problem(rng, 32, 32, M=900)
    perm = rng.permutation(900) + 1
    sign = np.where(rng.uniform(size=900) < 0.5, -1, 1)
    p.uvidxfcv = (perm[:200] * sign[:200]).astype(np.int32)
    p.uvidxamp = (perm[200:350] * sign[200:350]).astype(np.int32)
    with _mk(p, tune=dict(precision=precision)) as g:
        _check_cost(p, x, g)


@pytest.mark.parametrize("tune", [dict(bin=8, sub_points=64), dict(bin=32, sub_points=512, col_group=4),
                                  dict(chunk=64), dict(chisq_threads=512), dict(chisq_threads=512, chunk=64)])
def test_tunables_do_not_change_results(gpu_lib, tune):
    rng = np.random.default_rng(21)
    p, x = random_problem(rng, 40, 40, M=700)
    with _mk(p) as g, _mk(p, tune=tune) as h:
        c1, gr1, _ = g.cost_grad(x)
        c2, gr2, _ = h.cost_grad(x)
    assert abs(c1 - c2) <= 1e-6 * abs(c1)
    assert rel_l2(gr2, gr1) < 1e-6


@pytest.mark.parametrize("tune", [dict(chisq_threads=128), dict(chisq_threads=1024), dict(fft_radix=3),
                                  dict(chunk=8), dict(bin=12), dict(dft_shapes=4)])
def test_invalid_tunables_are_rejected(gpu_lib, tune):
    from smili_amd._lib import SmiliError
    rng = np.random.default_rng(22)
    p, _ = random_problem(rng, 16, 16, M=50)
    with pytest.raises(SmiliError) as e:
        _mk(p, tune=tune)
    assert e.value.code == -1 and next(iter(tune)) in str(e.value)


def test_points_on_the_period_boundary(gpu_lib):
    """uv points at exactly +-pi and beyond the principal period wrap
    (the transform is 2 pi periodic in u' for integer modes)."""
    rng = np.random.default_rng(3)
    p, x = random_problem(rng, 16, 16, M=50)
    p.u[:10] = np.pi
    p.u[10:20] = -np.pi
    p.v[20:30] = np.pi
    p.u[30:40] = 2.5 * np.pi
    with _mk(p) as g:
        V = g.model_vis(x)
    Vd = O.dft_fwd(O.image_1d_to_2d(x, p.xidx, p.yidx, 16, 16), p.u, p.v)
    assert rel_l2(V, Vd) <= TOL_V


def test_nonfinite_is_reported(gpu_lib):
    """A zero image gives |V| = 0 and the amplitude gradient divides by it
    (fftlib.f90:488): the reference returns a finite cost with a NaN
    gradient silently; the ABI returns the same values and flags them."""
    from smili_amd._lib import SmiliError
    rng = np.random.default_rng(4)
    p, x = random_problem(rng, 16, 16, M=50, nfcv=0, ncp=0, nca=0, regs=())
    with _mk(p) as g:
        with pytest.raises(SmiliError) as e:
            g.cost_grad(np.zeros_like(x))
        assert e.value.code == -4
        c, gr, t = g.cost_grad(np.zeros_like(x), allow_nonfinite=True)
    c_ref = O.calc_cost(np.zeros_like(x), p)[0]
    assert math.isfinite(c) and abs(c - c_ref) <= 1e-12 * abs(c_ref)
    assert not np.all(np.isfinite(gr))


def test_chunks_split_at_the_slot_limit(gpu_lib):
    """Random indices over 3000 uv points: every chi-square work item hits
    the 1024 distinct-point limit and splits."""
    rng = np.random.default_rng(41)
    p, x = random_problem(rng, 32, 32, M=3000, nfcv=500, namp=800, ncp=1500, nca=2000)
    with _mk(p) as g:
        _check_cost(p, x, g)


@pytest.mark.parametrize("shape,window", [((24, 20), False), ((33, 17), True), ((64, 64), False)])
def test_fp64_mode_matches_fp64_es_oracle(gpu_lib, shape, window):
    """precision=64 runs the same ES-kernel NUFFT in FP64: it must agree with
    the FP64 ES oracle to rounding (visibilities rel-l2 <= 1e-11, cost rel
    <= 1e-12, gradient rel-l2 <= 1e-10)."""
    rng = np.random.default_rng(300 + shape[0])
    p, x = random_problem(rng, shape[0], shape[1], M=500, window=window)
    I2d = O.image_1d_to_2d(x, p.xidx, p.yidx, p.Nx, p.Ny)
    with _mk(p, tune=dict(precision=64)) as g:
        V = g.model_vis(x)
        c, gr, t = g.cost_grad(x)
    Ve = O.es_nufft_type2(I2d, p.u, p.v)
    c_es, g_es, _ = O.calc_cost(x, p, fwd=O.es_nufft_type2, adj=O.es_nufft_type1)
    print("fp64", shape, "V", rel_l2(V, Ve), "cost", abs(c - c_es) / abs(c_es), "grad", rel_l2(gr, g_es))
    assert rel_l2(V, Ve) <= 1e-11
    assert abs(c - c_es) <= 1e-12 * abs(c_es)
    assert rel_l2(gr, g_es) <= 1e-10


def test_reported_chisq_terms_at_one_sigma(gpu_lib):
    """Data one standard deviation off the ES-kernel model: the GPU reports
    chisq_fcv = chisq_amp = chisq_cp = chisq_ca = 1 (the re-normalisation
    of fftlib.f90:342-366) on the FP64 path."""
    from tests.test_oracle_closed_forms import one_sigma_problem
    rng = np.random.default_rng(18)
    p, x = one_sigma_problem(rng, fwd=O.es_nufft_type2)
    with _mk(p, tune=dict(precision=64)) as g:
        c, gr, t = g.cost_grad(x)
    for k in ("chisq_fcv", "chisq_amp", "chisq_cp", "chisq_ca"):
        assert t[k] == pytest.approx(1.0, abs=1e-9), k


def test_host_buffers_give_the_same_evaluation(gpu_lib):
    """The handle's page-locked x / grad (what the imager's L-BFGS-B loop
    works in) against the caller's own arrays: the same cost, the gradient
    to the spreading atomics' rounding."""
    rng = np.random.default_rng(19)
    p, x = random_problem(rng, 32, 24, M=400, window=True)
    with _mk(p, tune=dict(precision=64)) as g:
        c1, g1, t1 = g.cost_grad(x)
        hx, hg = g.host_buffers()
        hx[:] = x
        c2, g2, t2 = g.cost_grad(hx, out=hg)
        assert g2 is hg
        g2 = g2.copy()
        # own arrays again after the pinned call: the staging does not leak
        c3, g3, _ = g.cost_grad(x * 1.01)
        c4, g4, _ = g.cost_grad(x)
    assert c1 == c2 == c4 and c3 != c1
    assert t1 == t2
    assert rel_l2(g2, g1) <= 1e-13 and rel_l2(g4, g1) <= 1e-13


def test_gradient_reproducible_at_full_size(gpu_lib):
    """smili_tune.deterministic at the 256^2 image with 1e5 amplitudes and
    the closure data of 12 stations x 500 timestamps (FP64 and FP32 NUFFT):
    repeated evaluations, through the host path and on device buffers, and
    graph replays give the same cost and gradient bit for bit (fixed-order
    sums throughout, no float atomics across waves); they equal the default
    path's to rounding."""
    import torch
    from smili_amd import workloads
    from smili_amd.cost import CostProblem
    prep, _ = workloads.single_large()
    for prec in (64, 32):
        with CostProblem(**prep.kwargs, tune=dict(precision=prec)) as f:
            cf, gf, _ = f.cost_grad(prep.x0)
        with CostProblem(**prep.kwargs, tune=dict(precision=prec, deterministic=1)) as g, \
                CostProblem(**prep.kwargs, tune=dict(precision=prec, deterministic=1, use_graph=1)) as h:
            c0, g0, t0 = g.cost_grad(prep.x0)
            assert abs(c0 - cf) <= (1e-13 if prec == 64 else 1e-6) * abs(cf)
            assert rel_l2(g0, gf) <= (1e-12 if prec == 64 else 1e-5)
            for _ in range(2):
                ch, gh, th = h.cost_grad(prep.x0)
                assert ch == c0 and th == t0 and np.array_equal(gh, g0)
            for _ in range(3):
                c1, g1, t1 = g.cost_grad(prep.x0)
                assert c1 == c0 and t1 == t0
                assert np.array_equal(g1, g0)
            xd = torch.from_numpy(prep.x0).to("cuda")
            gd = torch.empty_like(xd)
            for _ in range(3):
                g.cost_grad_device(xd.data_ptr(), gd.data_ptr())
                g.synchronize()
                assert np.array_equal(gd.cpu().numpy(), g0)


@pytest.mark.parametrize("nimg", [32, 33])
def test_host_path_on_both_sides_of_the_mapped_limit(gpu_lib, nimg):
    """The host-pointer entry point moves x and the gradient through the
    handle's mapped buffers up to 2^19 pixels (32 images of 128^2 exactly)
    and by DMA beyond (33): both give the device entry point's cost and
    terms bit for bit and its gradient to rounding."""
    import torch
    from smili_amd.cost import BatchProblem
    rng = np.random.default_rng(90 + nimg)
    cases = [random_problem(rng, 128, 128, M=120, nfcv=80, namp=20, ncp=10, nca=10, regs=("l1", "tsv"))
             for _ in range(nimg)]
    kws = [problem_kwargs(p) for p, _ in cases]
    X = np.concatenate([x for _, x in cases])
    assert (X.size <= 1 << 19) == (nimg == 32)
    with BatchProblem(kws) as b:
        costs, grad, terms = b.cost_grad(X)
        xd = torch.from_numpy(X).to("cuda")
        gd = torch.empty_like(xd)
        td = torch.empty(b.nprob * b.nterms, dtype=torch.float64, device="cuda")
        b.cost_grad_device(xd.data_ptr(), gd.data_ptr(), td.data_ptr())
        b.synchronize()
        # and the host path again, after the device one
        costs2, grad2, _ = b.cost_grad(X)
    tdev = td.cpu().numpy().reshape(b.nprob, b.nterms)
    assert np.array_equal(np.asarray(terms).reshape(b.nprob, b.nterms), tdev)
    assert np.array_equal(costs, costs2)
    assert rel_l2(gd.cpu().numpy(), grad) <= 1e-13 and rel_l2(grad2, grad) <= 1e-13
