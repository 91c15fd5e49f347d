"""Pins of the oracle's radial-node loop nests and k-space tables (CPU only).

Each C loop nest is checked against an independent numpy formulation on seeded
inputs (einsum / vectorised sums written from the mathematical definition, not
the loop order), and the k-space closed form F_L is checked against direct
numerical quadrature of its defining integral.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn

import oracle


@pytest.fixture
def rng():
    return np.random.default_rng(11)


def test_apply_fl_weights(rng):
    lmax, nr = 12, 5
    ls, _ = oracle.getlm(lmax)
    f = rng.standard_normal((lmax + 1, 3, nr))
    h = rng.standard_normal((2, len(ls))) + 1j * rng.standard_normal((2, len(ls)))
    out = oracle.apply_fl_weights(f, h, ls, 0.7)
    ref = 0.7 * (np.einsum('lr,l->rl', f[ls, 0], h[0]) + np.einsum('lr,l->rl', f[ls, 1], h[1]))
    assert np.allclose(out, ref, rtol=1e-14, atol=0)


def test_gnl_sums(rng):
    nr, npix = 4, 300
    w = rng.random(nr)
    P, Q, P2, Q2 = (rng.standard_normal((nr, npix)) for _ in range(4))
    assert np.isclose(oracle.gnl_loc_sum(w, P, Q), np.sum(w[:, None] * P ** 3 * Q), rtol=1e-13)
    ref = np.sum(w[:, None] * (P ** 2 * P2 * Q2 + P2 ** 2 * P * Q))
    assert np.isclose(oracle.gnl_loc_disc_sum(w, P, P2, Q, Q2), ref, rtol=1e-13)
    # the Monte-Carlo collapse used by the batched GPU path
    sims = [(rng.standard_normal((nr, npix)), rng.standard_normal((nr, npix))) for _ in range(3)]
    U = sum(a * b for a, b in sims)
    V = sum(a * a for a, _ in sims)
    coll = np.sum(w[:, None] * (P ** 2 * U + P * Q * V))
    loop = sum(oracle.gnl_loc_disc_sum(w, P, a, Q, b) for a, b in sims)
    assert np.isclose(coll, loop, rtol=1e-12)


def test_radial_sum_and_integrate_pq(rng):
    lmax, nr = 10, 6
    ls, _ = oracle.getlm(lmax)
    lm = rng.standard_normal((nr, len(ls))) + 1j * rng.standard_normal((nr, len(ls)))
    w = rng.random(nr)
    f = rng.standard_normal((lmax + 2, 3, nr))
    ref = np.einsum('ri,ixr,r->xi', lm, f[ls], w)
    assert np.allclose(oracle.radial_sum(lm, w, f, ls), ref, rtol=1e-13)
    out = np.zeros((3, len(ls)), dtype=complex)
    oracle.integrate_pq(f, np.ascontiguousarray(lm.T), ls, w, out)
    assert np.allclose(out, 0.5 * ref, rtol=1e-13)


def test_tau_sum_n0_and_convolve(rng):
    Lmax, nr, Lmin = 9, 4, 2
    Ls, Ms = oracle.getlm(Lmax)
    sel = Ls >= 1
    Ls, Ms = Ls[sel], Ms[sel]
    A = rng.standard_normal((1, len(Ls), nr)) + 1j * rng.standard_normal((1, len(Ls), nr))
    B = rng.standard_normal((1, len(Ls), nr)) + 1j * rng.standard_normal((1, len(Ls), nr))
    F = rng.standard_normal((Lmax + 1, nr, nr))
    w = rng.random(nr)
    w[2] = 0.
    tau = oracle.tau_sum_n0(A, B, F, w, Ls, Ms, Lmin, Lmax)
    mw = (1. + (Ms > 0)) * ((Ls >= Lmin) & (Ls <= Lmax))
    ref = np.einsum('i,ijk,j,k,ij,ik->', mw, F[Ls], w, w, A[0], B[0].conj()).real / np.sqrt(4 * np.pi)
    assert np.isclose(tau, ref, rtol=1e-12)
    Li = np.arange(len(Ls), dtype=np.int32)
    conv = oracle.convolve_F_n000(A[0], F, w, Ls, Li, len(Ls))
    assert np.allclose(conv, np.einsum('ijk,k,ik->ij', F[Ls], w, A[0]) / np.sqrt(4 * np.pi), rtol=1e-13)


def test_outer_product(rng):
    nin = 50
    Qa = rng.standard_normal((4, 2, nin)) + 1j * rng.standard_normal((4, 2, nin))
    Qb = rng.standard_normal((4, 3, nin)) + 1j * rng.standard_normal((4, 3, nin))
    M = np.array([[1, 0, 0, -3], [0, 1, -3, 0], [0, -3, 9, 0], [-3, 0, 0, 9]], float)
    ref = np.einsum('cd,caj,dbj->ab', M, Qa.conj(), Qb).real / (24 * 48)
    assert np.allclose(oracle.outer_product(Qa, Qb), ref, rtol=1e-13)


def test_q_integral_matches_definition():
    from polyspec_amd.k_integrals import q_integral
    k = np.geomspace(1e-4, 0.2, 80)
    r = np.array([13500., 14100.])
    lmax = 20
    T = np.zeros((1, lmax + 1, len(k)))
    for l in range(2, lmax + 1):
        T[0, l] = spherical_jn(l, k * 14000.)
    q = q_integral(k, T, r, 2, lmax, 1)
    l, ir = 7, 1
    f = 2 / np.pi * k ** 2 * T[0, l] * spherical_jn(l, k * r[ir])
    assert np.isclose(q[l, 0, ir], (-1) ** l * np.trapezoid(f, k), rtol=1e-12)
    assert np.all(q[:2] == 0)
