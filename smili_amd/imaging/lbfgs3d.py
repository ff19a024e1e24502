"""Dynamic imager (movies) with SMILI's API: ``imaging3d(initmovie, ...,
vistable=..., rt_lambda=..., ...)`` (smili/imaging/lbfgs3d.py:50-832).

The objective+gradient of fftim3d.calc_cost (src/fortlib/fftim3d.f90:
472-640) runs on the GPU (``MovieProblem``, the C ABI of libsmili_hip.so);
the optimiser is the same host FP64 L-BFGS-B as the static imager over the
Nt * Npix frame-major vector (fftim3d.f90:336-431: bounds x >= 0 when
nonneg, stop after niter + 1 iterations), with ``m``, ``factr``, ``pgtol``
from the module dict ``lbfgsbprms`` (lbfgs3d.py:34-38).
"""
from __future__ import annotations

import copy

import numpy as np

from ..cost import MovieProblem
from .lbfgs import _angconv, _window_of, solve
from .prep3d import prepare3d

lbfgsbprms = {
    "m": 5,
    "factr": 1e1,
    "pgtol": 0.,
}


def imaging3d(initmovie, imregion=None, vistable=None, amptable=None, bstable=None, catable=None,
              w_vis=1, w_amp=1, w_bs=1, w_ca=1,
              l1_lambda=-1., l1_prior=None, l1_noise=1e-2, l1_type=1, l1_floor=1e-1,
              sm_lambda=-1, sm_maj=50., sm_min=50., sm_phi=0.,
              tv_lambda=-1., tv_prior=None, tsv_lambda=-1., tsv_prior=None,
              kl_lambda=-1., kl_prior=None, gs_lambda=-1, gs_prior=None,
              tfd_lambda=-1, tfd_tgterror=0.01,
              lc_lambda=-1, lc_array=None, lc_tgterror=0.01, lc_normalize=False,
              cen_lambda=-1, cen_alpha=3,
              rt_lambda=-1., rt_prior=None, ri_lambda=-1., ri_prior=None,
              rs_lambda=-1, rs_prior=None, rf_lambda=-1, rf_prior=None,
              niter=1000, nonneg=True, nprint=500, totalflux=None, inorm=1.,
              istokes=0, ifreq=0, output='list', tune=None, verbose=True, return_info=False,
              warmstart=None):
    """smili.imaging.lbfgs3d.imaging3d on the GPU cost.

    ``initmovie`` has ``Nt`` and ``images[i]`` (IMFITS-like, or
    ``smili_amd.imaging.image.Movie``); tables carry a ``frmidx`` column.
    ``rs_*`` / ``rf_*`` are accepted and, as in the reference's cost
    (calc_cost_dynamical, src/fortlib/image3d.f90:280-392, which never
    evaluates them), have no effect; ``cen_*`` likewise has no effect on a
    movie.  Returns a copy of initmovie holding the solution
    (lbfgs3d.py:825-832); with ``return_info`` also the SolveInfo.

    ``warmstart``: a ``smili_amd.warmstart.unet.UNet`` handle (not in the
    reference).  Every frame is then proposed by the network from that
    frame's visibilities (``warmstart.movie.propose_movie``; the movie's
    pixel must be an integer multiple of the network's) and the solve
    starts from the proposal when its objective is below initmovie's;
    ``info.start`` says which start was used.
    """
    Nt = initmovie.Nt
    initimage = initmovie.images[0]
    hdr = initimage.header
    if warmstart is not None:
        # the proposals are made from each frame's visibilities and binned
        # onto a grid centred on its reference pixel with RA to the left
        if vistable is None:
            raise ValueError("warmstart needs vistable: the proposals come from the frames' visibilities")
        centred = (abs(hdr["nxref"] - (hdr["nx"] + 1) / 2.0) < 1e-9 and
                   abs(hdr["nyref"] - (hdr["ny"] + 1) / 2.0) < 1e-9)
        if not centred or not hdr["dx"] < 0:
            raise ValueError("warmstart needs the reference pixel at the image centre ((nx + 1) / 2, "
                             "(ny + 1) / 2) and dx < 0")
    window = _window_of(imregion, initimage, istokes, ifreq)

    def prior(p):
        return None if p is None else np.asarray(p.data)[0, 0]

    frames = np.stack([np.asarray(initmovie.images[i].data, np.float64)[istokes, ifreq] for i in range(Nt)])
    conv = np.deg2rad(_angconv(initimage))
    prep = prepare3d(frames, np.deg2rad(hdr["dx"]), np.deg2rad(hdr["dy"]), hdr["nxref"], hdr["nyref"], window,
                     vistable, amptable, bstable, catable, w_vis, w_amp, w_bs, w_ca,
                     l1_lambda, prior(l1_prior), l1_noise, l1_type, l1_floor,
                     sm_lambda, sm_maj * conv, sm_min * conv, sm_phi,
                     tv_lambda, prior(tv_prior), tsv_lambda, prior(tsv_prior),
                     kl_lambda, prior(kl_prior), gs_lambda, prior(gs_prior),
                     tfd_lambda, tfd_tgterror, lc_lambda, lc_array, lc_tgterror, lc_normalize,
                     cen_lambda, cen_alpha, rt_lambda, prior(rt_prior), ri_lambda, prior(ri_prior),
                     totalflux, inorm)
    with MovieProblem(**prep.kwargs, tune=tune) as prob:
        x0, start = prep.x0, "prior"
        if warmstart is not None:
            from ..warmstart.movie import propose_movie
            xw, _ = propose_movie(warmstart, vistable, prep, prob, abs(hdr["dx"]) * 3.6e9)
            if xw is not None and prob.cost_grad(xw)[0] < prob.cost_grad(prep.x0)[0]:
                x0, start = xw, "proposal"
            if verbose:
                print("Warm start: %s" % start)
        x, info = solve(prob, x0, niter=niter, nonneg=nonneg, nprint=nprint, m=lbfgsbprms["m"],
                        factr=lbfgsbprms["factr"], pgtol=lbfgsbprms["pgtol"],
                        inormfactr=prep.inormfactr, verbose=verbose)
        info.start = start
    Iout = x / prep.inormfactr
    out = copy.deepcopy(initmovie)
    Npix = prep.xidx.size
    for it in range(Nt):
        im = out.images[it]
        data = np.array(im.data, dtype=np.float64, copy=True)
        data[istokes, ifreq, prep.yidx - 1, prep.xidx - 1] = Iout[it * Npix:(it + 1) * Npix]
        im.data = data
        if hasattr(im, "update_fits"):
            im.update_fits()
    return (out, info) if return_info else out
