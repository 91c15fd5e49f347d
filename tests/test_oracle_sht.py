"""Pins of the FP64 oracle SHT (CPU only).

  * spin 0 and spin 2 synthesis / adjoint vs brute-force sums at every pixel
    (scipy sph_harm_y; Goldberg closed form) -- tests/golden/sht_bruteforce.npz;
  * odd spins (1, 3) vs the Goldberg brute force, and spin 1 vs the gradient of
    a spin-0 map (the HEALPix spin-1 convention);
  * the Goldberg spin-weighted harmonics used there vs spin-raising (eth)
    derivatives of scipy's Y_lm (pins the spin sign convention independently);
  * the Legendre recursion with exponent rescaling vs 60-digit mpmath values at
    l up to 2048 (tests/golden/ylm_spot.npz);
  * closed forms: a_00 -> 1/sqrt(4 pi), a_10 -> sqrt(3/4pi) cos(theta);
  * adjointness <Y a, m> = <a, Y^T m>_w and sum nphi = Npix.
"""
import os
import sys

import numpy as np
import pytest
from scipy.special import sph_harm_y

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

G = np.load(os.path.join(HERE, "golden", "sht_bruteforce.npz"))


def rel(a, b):
    return np.abs(np.asarray(a) - b).max() / np.abs(b).max()


@pytest.mark.parametrize("nside", [2, 4, 8])
def test_spin0_vs_bruteforce(nside):
    lmax = 3 * nside - 1
    a = G[f"s0_n{nside}_alm"]
    assert rel(oracle.synthesis(a[:, None], nside, lmax, 0)[:, 0], G[f"s0_n{nside}_map"]) < 1e-12
    m = G[f"s0_n{nside}_inmap"]
    assert rel(oracle.adjoint(m[:, None], nside, lmax, 0)[:, 0], G[f"s0_n{nside}_adj"]) < 1e-12


@pytest.mark.parametrize("nside", [2, 4])
def test_spin2_vs_bruteforce(nside):
    lmax = 3 * nside - 1
    GC = np.stack([G[f"s2_n{nside}_G"], G[f"s2_n{nside}_C"]], 1)
    RI = oracle.synthesis(GC, nside, lmax, 2)[0]
    assert rel(RI[0], G[f"s2_n{nside}_R"][0]) < 1e-12
    assert rel(RI[1], G[f"s2_n{nside}_I"][0]) < 1e-12
    adj = oracle.adjoint(np.stack([G[f"s2_n{nside}_Rin"], G[f"s2_n{nside}_Iin"]])[None], nside, lmax, 2)[0]
    assert rel(adj[0], G[f"s2_n{nside}_Gadj"]) < 1e-12
    assert rel(adj[1], G[f"s2_n{nside}_Cadj"]) < 1e-12


def _brute_spin(G_, C_, s, lmax, th, ph):
    """R + iI = -sum_{l, all m} (G + iC)_lm sY_lm with X_{l,-m} = (-1)^m conj(X_lm)."""
    ls, ms = oracle.getlm(lmax)
    out = np.zeros(len(th), complex)
    for i, (l, m) in enumerate(zip(ls, ms)):
        w = G_[i] + 1j * C_[i]
        out -= w * mg.goldberg_sylm(s, l, m, th, ph)
        if m > 0:
            wn = (-1) ** m * (np.conj(G_[i]) + 1j * np.conj(C_[i]))
            out -= wn * mg.goldberg_sylm(s, l, -m, th, ph)
    return out


@pytest.mark.parametrize("spin", [1, 3])
def test_odd_spin_vs_bruteforce(spin):
    """Odd spins pair +s with (-1)^s (-s) in the m < 0 half (sY_{l,-m} = (-1)^(s+m) conj(-sY_lm)),
    which even-spin tests cannot see; the reference's to_lm_spin carries the same (-1)^s
    (polyspec/base.py:255-256)."""
    nside, lmax = 4, 11
    th, ph = mg.healpix_centres(nside)
    a = mg.random_alm(np.random.default_rng(10 + spin), lmax, 2)
    ls, _ = oracle.getlm(lmax)
    a[:, ls < spin] = 0
    RI = oracle.synthesis(a[None], nside, lmax, spin)[0]
    want = _brute_spin(a[0], a[1], spin, lmax, th, ph)
    assert rel(RI[0] + 1j * RI[1], want) < 1e-12


def test_spin1_synthesis_is_the_gradient():
    """HEALPix spin-1 convention: G = sqrt(l(l+1)) a_lm, C = 0 synthesises
    (d_theta T, d_phi T / sin theta) of T = sum a_lm Y_lm (finite differences of scipy's Y_lm)."""
    nside, lmax = 4, 9
    th, ph = mg.healpix_centres(nside)
    a = mg.random_alm(np.random.default_rng(3), lmax, 1)[0]
    ls, ms = oracle.getlm(lmax)

    def T(t, p):
        out = np.zeros(len(t))
        for i, (l, m) in enumerate(zip(ls, ms)):
            y = sph_harm_y(l, m, t, p)
            out += (a[i] * y).real * (2 if m else 1)
        return out
    h = 1e-5
    dth = (T(th + h, ph) - T(th - h, ph)) / (2 * h)
    dph = (T(th, ph + h) - T(th, ph - h)) / (2 * h) / np.sin(th)
    GC = np.stack([a * np.sqrt(ls * (ls + 1.)), np.zeros_like(a)])[None]
    RI = oracle.synthesis(GC, nside, lmax, 1)[0]
    assert rel(RI[0], dth) < 1e-8
    assert rel(RI[1], dph) < 1e-8


def _eth(f, s, th, ph, h=1e-5):
    """eth f = -(sin th)^s (d_th + i/sin th d_ph) [(sin th)^-s f] by central differences."""
    g = lambda t, p: np.sin(t) ** (-s) * f(t, p)
    dth = (g(th + h, ph) - g(th - h, ph)) / (2 * h)
    dph = (g(th, ph + h) - g(th, ph - h)) / (2 * h)
    return -np.sin(th) ** s * (dth + 1j / np.sin(th) * dph)


@pytest.mark.parametrize("l,m", [(2, 0), (3, 1), (4, -2), (5, 3), (3, -3)])
def test_goldberg_matches_eth_of_scipy_ylm(l, m):
    """sY_lm = sqrt((l-s)!/(l+s)!) eth^s Y_lm (Goldberg et al. 1967) for s = 1, 2."""
    th = np.array([0.4, 1.1, 2.0])
    ph = np.array([0.3, 2.5, 4.0])
    y0 = lambda t, p: sph_harm_y(l, m, t, p)
    y1 = mg.goldberg_sylm(1, l, m, th, ph)
    e1 = _eth(y0, 0, th, ph) / np.sqrt(l * (l + 1.))
    assert np.abs(y1 - e1).max() < 1e-7
    y2 = mg.goldberg_sylm(2, l, m, th, ph)
    e2 = _eth(lambda t, p: mg.goldberg_sylm(1, l, m, t, p), 1, th, ph) / np.sqrt((l - 1.) * (l + 2.))
    assert np.abs(y2 - e2).max() < 1e-6
    # s = 0 limit is scipy's Y_lm
    assert np.abs(mg.goldberg_sylm(0, l, m, th, ph) - y0(th, ph)).max() < 1e-12


def test_recursion_spot_values_large_l():
    S = np.load(os.path.join(HERE, "golden", "ylm_spot.npz"))
    for (l, m, t), v in zip(S["spots"], S["vals"]):
        o = oracle.spin_legendre(int(l), int(m), 0, t)[int(l)]
        if abs(v) < 1e-100:
            assert abs(o) < 1e-100
        else:
            assert abs(o - v) < 1e-10 * abs(v), (l, m, t, o, v)


def test_closed_forms_and_geometry():
    nside, lmax = 8, 10
    th, _ = oracle.pix_angles(nside)
    ls, ms = oracle.getlm(lmax)
    a = np.zeros((1, 1, len(ls)), complex)
    a[0, 0, 0] = 1.0
    assert np.allclose(oracle.synthesis(a, nside, lmax)[0, 0], 1 / np.sqrt(4 * np.pi), atol=1e-14)
    a[0, 0, 0] = 0
    a[0, 0, 1] = 1.0
    assert np.allclose(oracle.synthesis(a, nside, lmax)[0, 0], np.sqrt(3 / (4 * np.pi)) * np.cos(th), atol=1e-13)
    ri = oracle.ring_info(nside)
    assert sum(r[3] for r in ri) == 12 * nside * nside
    starts = [r[4] for r in ri]
    assert starts == sorted(starts) and starts[0] == 0


@pytest.mark.parametrize("spin", [0, 1, 2])
def test_oracle_adjointness(spin):
    nside, lmax = 16, 40
    rng = np.random.default_rng(spin)
    ls, ms = oracle.getlm(lmax)
    nc = 2 if spin else 1
    a = rng.standard_normal((1, nc, len(ls))) + 1j * rng.standard_normal((1, nc, len(ls)))
    a[..., ms == 0] = a[..., ms == 0].real
    a[..., ls < spin] = 0
    m = rng.standard_normal((1, nc, 12 * nside ** 2))
    lhs = np.sum(oracle.synthesis(a, nside, lmax, spin) * m)
    YTm = oracle.adjoint(m, nside, lmax, spin)
    rhs = np.sum((np.conj(a) * YTm).real * (1. + (ms > 0)))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs) + 1e-10
