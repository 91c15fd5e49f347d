"""Mask builders and RING geometry (polyspec_amd/masks.py) on the CPU."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import healpix_centres  # noqa: E402  (geometry written independently)

from polyspec_amd.masks import galactic_cut_mask, pixel_z_phi


def test_ring_geometry_matches_independent_restatement():
    for nside in (1, 2, 4, 16, 64):
        z, ph = pixel_z_phi(nside)
        th, ph2 = healpix_centres(nside)
        assert np.allclose(z, np.cos(th), atol=1e-14)
        assert np.allclose(np.mod(ph - ph2 + np.pi, 2 * np.pi) - np.pi, 0, atol=1e-12)


def test_galactic_cut_mask_structure():
    nside, cut, hole, apo = 256, 20., 0.5, 1.
    m = galactic_cut_mask(nside, cut_deg=cut, nholes=40, hole_deg=hole, apo_deg=apo, seed=3)
    z, ph = pixel_z_phi(nside)
    b = np.degrees(np.arcsin(z))
    assert m.shape == (12 * nside ** 2,) and m.min() >= 0 and m.max() <= 1
    assert np.all(m[np.abs(b) < cut] == 0)
    # the apodised band rises monotonically in |b| away from holes
    nohole = galactic_cut_mask(nside, cut_deg=cut, nholes=0, hole_deg=hole, apo_deg=apo)
    band = (np.abs(b) > cut) & (np.abs(b) < cut + apo)
    assert np.all((nohole[band] > 0) & (nohole[band] < 1))
    assert np.all(nohole[np.abs(b) >= cut + apo] == 1)
    mid = np.abs(np.abs(b) - (cut + apo / 2)) < 0.05
    assert np.allclose(nohole[mid], 0.5, atol=0.1)
    # every hole centre outside the cut is zeroed; the holes remove about their area
    rng = np.random.default_rng(3)
    zc, pc = rng.uniform(-1, 1, 40), rng.uniform(0, 2 * np.pi, 40)
    for zi, pi in zip(zc, pc):
        if abs(np.degrees(np.arcsin(zi))) > cut + hole + apo:
            cosd = z * zi + np.sqrt(1 - z * z) * np.sqrt(1 - zi * zi) * np.cos(ph - pi)
            assert m[np.argmax(cosd)] == 0
    lost = np.sum(nohole - m) / len(m)
    area = 40 * np.pi * np.radians(hole + apo / 2) ** 2 / (4 * np.pi) * np.mean(nohole)
    assert 0.5 * area < lost < 1.5 * area
