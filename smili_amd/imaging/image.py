"""Light image container with SMILI's IMFITS conventions.

``header`` holds nx, ny, dx, dy (degrees; dx < 0 for RA increasing to the
left), nxref, nyref (1-based reference pixel, default (n+1)/2) as in
smili/imdata/imdata.py:93-97; ``data`` is [nstokes, nfreq, ny, nx]
(imdata.py:39).  Pixel (ix, iy) sits at x = dx (ix - nxref),
y = dy (iy - nyref) (imdata.py:726-747).
"""
from __future__ import annotations

import copy

import numpy as np

_DEG_PER = {"deg": 1.0, "degree": 1.0, "arcmin": 1 / 60, "amin": 1 / 60, "arcsec": 1 / 3600,
            "asec": 1 / 3600, "mas": 1 / 3.6e6, "uas": 1 / 3.6e9, "rad": 180 / np.pi}


class Image:
    def __init__(self, nx, ny, dx, dy=None, angunit="uas", nxref=None, nyref=None, data=None):
        """dx, dy in `angunit`; dx is the RA increment and is stored negative
        (east to the left) unless given negative already."""
        self.angunit = angunit
        c = _DEG_PER[angunit]
        dy = abs(dx) if dy is None else dy
        self.header = dict(nx=int(nx), ny=int(ny), dx=-abs(dx) * c, dy=dy * c,
                           nxref=(nx + 1) / 2.0 if nxref is None else float(nxref),
                           nyref=(ny + 1) / 2.0 if nyref is None else float(nyref))
        self.data = np.zeros((1, 1, ny, nx)) if data is None else np.asarray(data, np.float64).reshape(1, 1, ny, nx)

    @property
    def dx_rad(self):
        return np.deg2rad(self.header["dx"])

    @property
    def dy_rad(self):
        return np.deg2rad(self.header["dy"])

    def get_xygrid(self, twodim=False, angunit=None):
        """imdata.py:726-747."""
        c = 1.0 / _DEG_PER[angunit or self.angunit]
        h = self.header
        xg = h["dx"] * (np.arange(h["nx"]) - h["nxref"] + 1) * c
        yg = h["dy"] * (np.arange(h["ny"]) - h["nyref"] + 1) * c
        return np.meshgrid(xg, yg) if twodim else (xg, yg)

    def copy(self):
        return copy.deepcopy(self)

    def totalflux(self, istokes=0, ifreq=0):
        return float(np.sum(self.data[istokes, ifreq]))


def nrmse(image, ref, istokes=0, ifreq=0):
    """NRMSE = sqrt(sum (I - K)^2 / sum K^2) (imdata.py:2311-2375, on images
    of the same grid)."""
    a = np.asarray(getattr(image, "data", image), np.float64)
    b = np.asarray(getattr(ref, "data", ref), np.float64)
    if a.ndim == 4:
        a = a[istokes, ifreq]
    if b.ndim == 4:
        b = b[istokes, ifreq]
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


class Movie:
    """Frames of one geometry with the attributes imaging3d reads from
    SMILI's MOVIE (smili/imdata/movie.py:46-90: ``Nt``, ``images[i]`` IMFITS-like
    with ``header`` / ``data``)."""

    def __init__(self, images):
        self.images = list(images)

    @classmethod
    def from_frames(cls, frames, dx, dy=None, angunit="uas", nxref=None, nyref=None):
        frames = np.asarray(frames, np.float64)
        Nt, ny, nx = frames.shape
        return cls([Image(nx, ny, dx, dy, angunit, nxref, nyref, frames[i]) for i in range(Nt)])

    @property
    def Nt(self):
        return len(self.images)

    def frames(self, istokes=0, ifreq=0):
        return np.stack([np.asarray(im.data)[istokes, ifreq] for im in self.images])

    def copy(self):
        return copy.deepcopy(self)
