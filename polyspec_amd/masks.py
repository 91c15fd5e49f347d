"""Synthetic sky masks and HEALPix RING geometry helpers (host, FP64).

The masks are the inputs the masked-sky analyses take (TSpecTemplate(mask=...),
Weightings(mask=...), the `mask` argument of polyspec/tspec_template.py:43 and
weighting.py:6): a galactic-latitude cut plus circular point-source holes, each
edge apodised with a raised cosine.  Pixel centres follow the HEALPix RING scheme
(Gorski et al. 2005, sec. 4).
"""
import numpy as np


def pixel_z_phi(nside):
    """z = cos(theta) and phi of every RING pixel, as two (12 nside^2,) arrays."""
    N = int(nside)
    npix = 12 * N * N
    z = np.empty(npix)
    ph = np.empty(npix)
    p = 0
    for i in range(1, 4 * N):
        ir = min(i, 4 * N - i)
        if ir < N:  # polar cap rings: 4 ir pixels
            zz, n, p0 = 1 - ir * ir / (3. * N * N), 4 * ir, np.pi / (4. * ir)
        else:  # equatorial belt: 4 N pixels, every other ring shifted by half a pixel
            zz, n = 2. * (2 * N - ir) / (3. * N), 4 * N
            p0 = np.pi / (4. * N) if ((ir - N) & 1) == 0 else 0.
        if i > 2 * N:
            zz = -zz
        z[p:p + n] = zz
        ph[p:p + n] = p0 + 2 * np.pi * np.arange(n) / n
        p += n
    return z, ph


def _raised_cosine(d, width):
    """0 for d <= 0, 1 for d >= width, 0.5 (1 - cos(pi d / width)) in between."""
    if width <= 0:
        return (d > 0).astype(np.float64)
    return np.where(d <= 0, 0., 0.5 * (1 - np.cos(np.pi * np.clip(d / width, 0, 1))))


def galactic_cut_mask(nside, cut_deg=20., nholes=200, hole_deg=0.5, apo_deg=1., seed=1):
    """|b| < cut_deg removed (b the latitude 90 deg - theta), plus `nholes` holes of radius
    hole_deg centred on seeded uniform points of the sphere; every edge apodised by a raised
    cosine over apo_deg (the mask reaches 1 at apo_deg from an edge).  Holes multiply."""
    z, ph = pixel_z_phi(nside)
    b = np.degrees(np.arcsin(np.clip(z, -1, 1)))
    m = _raised_cosine(np.abs(b) - cut_deg, apo_deg)
    rng = np.random.default_rng(seed)
    zc = rng.uniform(-1, 1, nholes)
    pc = rng.uniform(0, 2 * np.pi, nholes)
    st = np.sqrt(1 - z * z)
    rad = np.radians(hole_deg + apo_deg)
    negz = -z  # RING order: z decreases with the pixel index
    for zi, pi in zip(zc, pc):
        th = np.arccos(zi)
        # the rings within the hole's reach
        lo = np.searchsorted(negz, -np.cos(max(th - rad, 0.)), side='left')
        hi = np.searchsorted(negz, -np.cos(min(th + rad, np.pi)), side='right')
        sl = slice(lo, hi)
        cosd = z[sl] * zi + st[sl] * np.sqrt(1 - zi * zi) * np.cos(ph[sl] - pi)
        near = cosd > np.cos(rad)
        dd = np.degrees(np.arccos(np.clip(cosd[near], -1, 1))) - hole_deg
        m[np.arange(lo, hi)[near]] *= _raised_cosine(dd, apo_deg)
    return m
