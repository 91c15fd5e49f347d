"""HEALPix RING geometry (polyspec_amd/healpix.py) against what the pixelisation fixes on its own:
pixel centres against the ring formulas the SHTs use (masks.pixel_z_phi, pinned with the SHT
against brute-force sums at the centres), RING <-> NEST as inverse permutations with the NEST
hierarchy (the 4 children of a pixel surround its centre), and the 8-neighbour table against
brute force: two pixels are neighbours exactly when they share a vertex, the vertices computed
from the face coordinates.  Plus the order of healpy's table: E and W on the pixel's own ring,
N above, S below."""
import numpy as np
import pytest

from polyspec_amd import healpix as H
from polyspec_amd.masks import pixel_z_phi


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 32])
def test_pixel_centres_match_ring_formulas(nside):
    z, phi = pixel_z_phi(nside)
    th, ph = H.pix2ang(nside, np.arange(12 * nside ** 2))
    assert np.abs(np.cos(th) - z).max() < 1e-13
    d = np.abs(np.mod(ph - phi + np.pi, 2 * np.pi) - np.pi)
    assert d.max() < 1e-12


@pytest.mark.parametrize("nside", [1, 2, 4, 16])
def test_ring_nest_permutations_and_hierarchy(nside):
    npix = 12 * nside ** 2
    pix = np.arange(npix)
    nest = H.ring2nest(nside, pix)
    assert np.array_equal(np.sort(nest), pix)
    assert np.array_equal(H.nest2ring(nside, nest), pix)
    if nside > 1:
        # children 4p .. 4p+3 (NEST, nside) of pixel p (NEST, nside / 2) lie around its centre
        def vec(ns, ring_pix):
            th, ph = H.pix2ang(ns, ring_pix)
            return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
        par = vec(nside // 2, H.nest2ring(nside // 2, np.arange(npix // 4)))
        ch = vec(nside, H.nest2ring(nside, np.arange(npix))).reshape(-1, 4, 3)
        cos = np.einsum('pcx,px->pc', ch, par)
        size = np.sqrt(4 * np.pi / (npix // 4))
        assert np.arccos(np.clip(cos, -1, 1)).max() < 0.8 * size
        m = ch.mean(1)
        m /= np.linalg.norm(m, axis=1)[:, None]
        assert np.arccos(np.clip(np.einsum('px,px->p', m, par), -1, 1)).max() < 0.2 * size


def _vertex_neighbours(nside):
    """Pixels sharing a vertex with each pixel (brute force over the pixel corners)."""
    n = nside
    f, ix, iy = H.ring2xyf(n)
    owners = {}
    for cx, cy in ((0, 0), (1, 0), (0, 1), (1, 1)):
        z, phi = H.face_point(f, (ix + cx) / n, (iy + cy) / n)
        s = np.sqrt(np.clip(1 - z * z, 0, None))
        v = np.round(np.stack([s * np.cos(phi), s * np.sin(phi), z], -1), 9) + 0.
        for p in range(12 * n * n):
            owners.setdefault(tuple(v[p]), set()).add(p)
    nb = [set() for _ in range(12 * n * n)]
    for ps in owners.values():
        for p in ps:
            nb[p] |= ps
    for p in range(12 * n * n):
        nb[p].discard(p)
    return nb


@pytest.mark.parametrize("nside", [1, 2, 4, 8])
def test_neighbours_are_the_vertex_sharing_pixels(nside):
    npix = 12 * nside ** 2
    nb = H.get_all_neighbours(nside, np.arange(npix))
    brute = _vertex_neighbours(nside)
    for p in range(npix):
        got = [q for q in nb[:, p] if q >= 0]
        assert len(got) == len(set(got)) and set(got) == brute[p], (p, got, brute[p])
    # seven neighbours at the three-face vertices (z = +-2/3): 8 vertices x 3 pixels
    assert np.sum(nb < 0) == 24
    # symmetric relation
    for p in range(npix):
        for q in nb[:, p]:
            if q >= 0:
                assert p in nb[:, q]


@pytest.mark.parametrize("nside", [2, 8, 32])
def test_neighbour_directions(nside):
    npix = 12 * nside ** 2
    z, _ = pixel_z_phi(nside)
    face = H.ring2xyf(nside)[0]
    nb = H.get_all_neighbours(nside, np.arange(npix))
    # directions are steps in the pixel's own face frame; checked where the step stays in the face
    same = lambda d: (nb[d] >= 0) & (face[np.maximum(nb[d], 0)] == face)
    for d in (1, 5):  # W, E: same ring
        ok = same(d)
        assert ok.sum() > 0 and np.abs(z[nb[d][ok]] - z[ok]).max() < 1e-14
    ok = same(3)  # N
    assert np.all(z[nb[3][ok]] > z[ok])
    ok = same(7)  # S
    assert np.all(z[nb[7][ok]] < z[ok])


def test_stencil_reverse_table():
    nbr, rev = H.stencil9(4)
    assert np.array_equal(nbr[:, 0], np.arange(nbr.shape[0]))
    for p in range(nbr.shape[0]):
        slots = [s for s in rev[p] if s >= 0]
        assert all(nbr.reshape(-1)[s] == p for s in slots)
        assert len(slots) == np.sum(nbr == p)
