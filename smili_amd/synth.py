"""Seeded synthetic imaging workloads.

Skies are analytic (Gaussian mixes, Gaussian-blurred rings), so their
visibilities are exact closed forms and need no transform:

* circular Gaussian of flux F, FWHM t at (x0, y0):
  V(u, v) = F exp(-2 pi^2 s^2 (u^2 + v^2)) exp(+2 pi i (u x0 + v y0)),
  s = t / (2 sqrt(2 ln 2));
* thin ring of diameter d:  V = F J0(pi d rho),  rho = |(u, v)|; blurring by a
  Gaussian multiplies by the Gaussian's envelope;
* uniform disk of diameter d:  V = F 2 J1(pi d rho) / (pi d rho);
* crescent: a disk minus a smaller disk of the same surface brightness
  shifted inside it (a disk's transform times exp(+2 pi i (u sx + v sy))
  for the shift).

The +2 pi i sign is SMILI's (iflag=0 in src/wrapfinufft/wrapfinufft.f90:108-111)
with sky positions x = dx (ix - Nxref) (smili/imdata/imdata.py:743-744).

uv coverage by Earth-rotation synthesis follows calc_uvw
(src/fortlib/coord.f90:37-59).
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass

import numpy as np
from scipy.special import j0, j1

UAS = math.pi / 180.0 / 3600.0 / 1e6      # 1 micro-arcsecond in radians
C_LIGHT = 299792458.0
R_EARTH = 6371.0e3


# ----------------------------------------------------------------------------
# skies
# ----------------------------------------------------------------------------
@dataclass
class Gaussian:
    flux: float
    fwhm: float      # rad
    x0: float = 0.0  # rad
    y0: float = 0.0  # rad


@dataclass
class Ring:
    flux: float
    diameter: float  # rad
    blur_fwhm: float  # rad


@dataclass
class Disk:
    flux: float
    diameter: float   # rad
    blur_fwhm: float = 0.0
    x0: float = 0.0
    y0: float = 0.0


@dataclass
class Crescent:
    """Disk of diameter d_out minus a disk of diameter d_in < d_out shifted
    by (sx, sy) (|shift| <= (d_out - d_in) / 2 keeps it inside), equal
    surface brightness; total flux `flux`."""
    flux: float
    d_out: float
    d_in: float
    sx: float
    sy: float
    blur_fwhm: float = 0.0
    x0: float = 0.0
    y0: float = 0.0


def _disk_env(d, rho):
    a = math.pi * d * rho
    out = np.ones_like(a)
    nz = a > 1e-12
    out[nz] = 2.0 * j1(a[nz]) / a[nz]
    return out


def _sig(fwhm):
    return fwhm / (2.0 * math.sqrt(2.0 * math.log(2.0)))


def sky_vis(sky, u, v):
    """Exact visibilities of a list of components at (u, v) in wavelengths."""
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    out = np.zeros(u.shape, np.complex128)
    rho2 = u * u + v * v
    for c in sky:
        if isinstance(c, Gaussian):
            s = _sig(c.fwhm)
            out += c.flux * np.exp(-2 * math.pi ** 2 * s * s * rho2) * \
                np.exp(2j * math.pi * (u * c.x0 + v * c.y0))
        elif isinstance(c, (Disk, Crescent)):
            rho = np.sqrt(rho2)
            s = _sig(c.blur_fwhm)
            env = np.exp(-2 * math.pi ** 2 * s * s * rho2) * np.exp(2j * math.pi * (u * c.x0 + v * c.y0))
            if isinstance(c, Disk):
                out += c.flux * _disk_env(c.diameter, rho) * env
            else:
                q = (c.d_in / c.d_out) ** 2
                f_out = c.flux / (1.0 - q)
                out += (f_out * _disk_env(c.d_out, rho)
                        - f_out * q * _disk_env(c.d_in, rho) * np.exp(2j * math.pi * (u * c.sx + v * c.sy))) * env
        else:
            s = _sig(c.blur_fwhm)
            out += c.flux * j0(math.pi * c.diameter * np.sqrt(rho2)) * \
                np.exp(-2 * math.pi ** 2 * s * s * rho2)
    return out


def sky_image(sky, Nx, Ny, dx, dy, Nxref=None, Nyref=None, oversample=4):
    """Pixelised image [Ny, Nx] (Jy/pixel) of the sky, pixel (ix, iy) centred
    at x = dx (ix - Nxref), y = dy (iy - Nyref), 1-based."""
    Nxref = (Nx + 1) / 2.0 if Nxref is None else Nxref
    Nyref = (Ny + 1) / 2.0 if Nyref is None else Nyref
    os_ = oversample
    sub = (np.arange(os_) + 0.5) / os_ - 0.5
    xs = (np.arange(1, Nx + 1)[:, None] - Nxref + sub[None, :]).reshape(-1) * dx
    ys = (np.arange(1, Ny + 1)[:, None] - Nyref + sub[None, :]).reshape(-1) * dy
    X, Y = np.meshgrid(xs, ys)
    img = np.zeros_like(X)
    for c in sky:
        if isinstance(c, Gaussian):
            s = _sig(c.fwhm)
            img += c.flux / (2 * math.pi * s * s) * \
                np.exp(-((X - c.x0) ** 2 + (Y - c.y0) ** 2) / (2 * s * s))
        elif isinstance(c, (Disk, Crescent)):
            raise NotImplementedError("sky_image renders Gaussians and rings only")
        else:
            # ring as an annulus profile: radial Gaussian of the blur width
            s = _sig(c.blur_fwhm)
            r = np.sqrt(X * X + Y * Y)
            prof = np.exp(-(r - c.diameter / 2) ** 2 / (2 * s * s))
            img += prof * c.flux / max(prof.sum() * abs(dx * dy) / os_ ** 2, 1e-300)
    img *= abs(dx * dy) / os_ ** 2
    return img.reshape(Ny, os_, Nx, os_).sum(axis=(1, 3))


def asymmetric_ring(flux, diameter, width_fwhm, angle, asym=0.7, n_beads=48):
    """A ring of diameter d as n_beads Gaussian beads of FWHM width_fwhm
    whose brightness follows 1 + asym cos(theta - angle) (theta measured
    from +y towards +x): a crescent whose bright side points at `angle`."""
    th = 2 * math.pi * np.arange(n_beads) / n_beads
    w = 1.0 + asym * np.cos(th - angle)
    w *= flux / w.sum()
    r = diameter / 2
    return [Gaussian(float(w[k]), width_fwhm, float(r * math.sin(th[k])), float(r * math.cos(th[k])))
            for k in range(n_beads)]


def random_gaussians(rng, n_min=3, n_max=6, total_flux=1.0, fov=200 * UAS,
                     fwhm_range=(10 * UAS, 40 * UAS)):
    n = int(rng.integers(n_min, n_max + 1))
    fl = rng.uniform(0.2, 1.0, n)
    fl *= total_flux / fl.sum()
    out = []
    for i in range(n):
        out.append(Gaussian(flux=float(fl[i]),
                            fwhm=float(rng.uniform(*fwhm_range)),
                            x0=float(rng.uniform(-fov / 2, fov / 2)),
                            y0=float(rng.uniform(-fov / 2, fov / 2))))
    return out


# ----------------------------------------------------------------------------
# uv coverage
# ----------------------------------------------------------------------------
def station_xyz(rng, n, lon_deg=(-180.0, 180.0), lat_deg=(-64.0, 64.0)):
    """Seeded random geodetic station positions on a spherical Earth
    (uniform in area within the longitude / latitude bands)."""
    s0, s1 = (math.sin(math.radians(a)) for a in lat_deg)
    lat = np.arcsin(rng.uniform(s0, s1, n))
    lon = np.radians(rng.uniform(lon_deg[0], lon_deg[1], n))
    return np.stack([R_EARTH * np.cos(lat) * np.cos(lon),
                     R_EARTH * np.cos(lat) * np.sin(lon),
                     R_EARTH * np.sin(lat)], axis=1)


def calc_uvw(gst_hour, ra, dec, bx, by, bz):
    """coord.f90:37-59 (baseline XYZ in wavelengths -> u, v, w)."""
    h = gst_hour * math.pi / 12.0 - ra
    cg, sg = np.cos(h), np.sin(h)
    cd, sd = math.cos(dec), math.sin(dec)
    blx = cg * bx - sg * by
    bly = sg * bx + cg * by
    u = bly
    v = -blx * sd + bz * cd
    w = blx * cd + bz * sd
    return u, v, w


def elevation_ok(xyz, gst_hour, ra, dec, cut_deg):
    """Source elevation above cut at each station (spherical Earth)."""
    h = gst_hour * math.pi / 12.0 - ra
    src = np.array([math.cos(dec) * math.cos(-h), math.cos(dec) * math.sin(-h), math.sin(dec)])
    up = xyz / np.linalg.norm(xyz, axis=1, keepdims=True)
    return up @ src > math.sin(math.radians(cut_deg))


@dataclass
class Scan:
    gst: float
    stations: np.ndarray           # visible station ids
    bl: list                       # [(i, j)] i < j
    u: np.ndarray                  # wavelengths
    v: np.ndarray


def earth_rotation_scans(rng, n_st, dec_deg, freq_hz, n_scans, hours,
                         elev_cut_deg=None, ra=0.0, gst0=0.0,
                         lon_deg=(-180.0, 180.0), lat_deg=(-64.0, 64.0)):
    """Scans at n_scans GSTs in [gst0, gst0 + hours) of n_st seeded stations;
    with elev_cut_deg only stations seeing the source above it take part."""
    xyz = station_xyz(rng, n_st, lon_deg, lat_deg) * freq_hz / C_LIGHT
    dec = math.radians(dec_deg)
    out = []
    for t in np.linspace(gst0, gst0 + hours, n_scans, endpoint=False):
        gst = float(t)
        if elev_cut_deg is None:
            vis = np.arange(n_st)
        else:
            vis = np.nonzero(elevation_ok(xyz, gst, ra, dec, elev_cut_deg))[0]
        bl = [(int(a), int(b)) for a, b in itertools.combinations(vis, 2)]
        if not bl:
            continue
        a = np.array([p[0] for p in bl])
        b = np.array([p[1] for p in bl])
        d = xyz[b] - xyz[a]
        u, v, _ = calc_uvw(gst, ra, dec, d[:, 0], d[:, 1], d[:, 2])
        out.append(Scan(gst, vis, bl, u, v))
    return out


# ----------------------------------------------------------------------------
# data tables (columns follow the reference CSV tables,
# docs/examples/uvdata/{visibility,cphase,logcamp}.csv)
# ----------------------------------------------------------------------------
def _complex_noise(rng, sigma, n):
    return sigma * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def vis_table(rng, sky, u, v, sigma):
    """Full complex visibility table: u, v, amp, phase (deg), sigma."""
    V = sky_vis(sky, u, v) + _complex_noise(rng, sigma, len(u))
    return dict(u=np.asarray(u, np.float64), v=np.asarray(v, np.float64),
                amp=np.abs(V), phase=np.degrees(np.angle(V)),
                sigma=np.full(len(u), float(sigma)))


def closure_tables(rng, sky, scans, sigma, n_ca_per_quad=2, with_vis=False):
    """Closure phases over all triangles and log closure amplitudes over all
    quadrangles (n_ca_per_quad of the 3 per quadrangle) of every scan.

    Closures are formed from per-baseline noisy visibilities, so closures of
    one scan share noise as they do in real data.  With ``with_vis`` the
    per-baseline visibility table is returned too.
    """
    gu, gv, keys = [], [], {}
    tri, quad = [], []
    for s, sc in enumerate(scans):
        base = len(gu)
        for k, bl in enumerate(sc.bl):
            keys[(s,) + bl] = base + k
        gu.extend(sc.u.tolist())
        gv.extend(sc.v.tolist())

        def ref(a, b):
            return (keys[(s, a, b)], 1) if a < b else (keys[(s, b, a)], -1)

        st = [int(x) for x in sc.stations]
        for a, b, c in itertools.combinations(st, 3):
            tri.append(ref(a, b) + ref(b, c) + ref(c, a))
        for a, b, c, d in itertools.combinations(st, 4):
            sets = [((a, b), (c, d), (a, c), (b, d)), ((a, d), (b, c), (a, c), (b, d)),
                    ((a, b), (c, d), (a, d), (b, c))][:n_ca_per_quad]
            for q in sets:
                quad.append(sum((ref(*p) for p in q), ()))
    gu = np.asarray(gu, np.float64)
    gv = np.asarray(gv, np.float64)
    Vt = sky_vis(sky, gu, gv)
    Vn = Vt + _complex_noise(rng, sigma, gu.size)

    def pick(arr, ncol):
        arr = np.asarray(arr, np.int64).reshape(-1, 2 * ncol)
        g = arr[:, 0::2]
        s = arr[:, 1::2].astype(np.float64)
        Vk = np.where(s > 0, Vn[g], np.conj(Vn[g]))
        Tk = Vt[g]
        return g, s, Vk, Tk

    cp, ca = {}, {}
    if tri:
        g, s, Vk, Tk = pick(tri, 3)
        B = Vk[:, 0] * Vk[:, 1] * Vk[:, 2]
        rel = np.sqrt(np.sum((sigma / np.maximum(np.abs(Tk), 1e-12)) ** 2, axis=1))
        for k, nm in enumerate(("12", "23", "31")):
            cp["u" + nm] = gu[g[:, k]] * s[:, k]
            cp["v" + nm] = gv[g[:, k]] * s[:, k]
        cp["amp"] = np.abs(B)
        cp["phase"] = np.degrees(np.angle(B))
        cp["sigma"] = np.abs(B) * rel
    if quad:
        g, s, Vk, Tk = pick(quad, 4)
        la = (np.log(np.abs(Vk[:, 0])) + np.log(np.abs(Vk[:, 1])) -
              np.log(np.abs(Vk[:, 2])) - np.log(np.abs(Vk[:, 3])))
        for k in range(4):
            ca["u%d" % (k + 1)] = gu[g[:, k]] * s[:, k]
            ca["v%d" % (k + 1)] = gv[g[:, k]] * s[:, k]
        ca["logamp"] = la
        ca["logsigma"] = np.sqrt(np.sum((sigma / np.maximum(np.abs(Tk), 1e-12)) ** 2, axis=1))
    if with_vis:
        vis = dict(u=gu, v=gv, amp=np.abs(Vn), phase=np.degrees(np.angle(Vn)),
                   sigma=np.full(gu.size, float(sigma)))
        return cp, ca, vis
    return cp, ca


def disk_uv(rng, n, radius):
    r = radius * np.sqrt(rng.uniform(0, 1, n))
    t = rng.uniform(0, 2 * math.pi, n)
    return r * np.cos(t), r * np.sin(t)
