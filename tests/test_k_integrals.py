"""Pins of the k-space precompute (polyspec_amd/k_integrals.py, restating
k_integrals.pyx:262-352, 453-494) by mathematics that does not go through its formulas.

  * F_L(r, r') = (2/pi) int_0^inf dk k^2 P_zeta(k) j_L(k r) j_L(k r')  (the quantity the
    reference's closed form evaluates, k_integrals.pyx:453-494 with the prefactor of
    tspec_template.py:663), computed by Gauss-Legendre panels, graded towards k = 0, up to
    K = 2e4 / r, plus the analytic non-oscillating tail (C / (pi r^2)) K^(ns-3) / (3 - ns) of
    the diagonal (j_L(x)^2 -> 1 / (2 x^2) on average); agreement to 1e-9 for L <= 10.
  * Scaling: P_zeta ~ k^(ns-4) gives F_L(lam r, lam r') = lam^(1-ns) F_L(r, r'), which carries the
    pin to the radii of the cosmological radial grids.
  * p_l and q_l (k_integrals.pyx:262-352) are trapezoid sums on the caller's k grid; with the
    transfer T_l(k) = j_l(k r0) (p) or P_zeta(k) j_l(k r0) (q) both integrals become
    (-1)^l F_l(r0, r), so a dense grid must reproduce the closed form (to 3e-7 here).  This catches
    a dropped (-1)^l, 2/pi or P_zeta, or an l / r index mix-up.
"""
import numpy as np
import pytest
from scipy.special import roots_legendre, spherical_jn

from polyspec_amd.k_integrals import F_integral_n0, Pzeta, p_integral_n0, q_integral

AS, NS, KP = 2.1e-9, 0.96, 0.05


def _quad_F(L, a, b, K=2e4, h=0.5, npt=20):
    x, w = roots_legendre(npt)
    edges = np.concatenate([[0.], h * 2.0 ** -np.arange(60, 0, -1), np.arange(h, K + h / 2, h)])
    lo, hi = edges[:-1], edges[1:]
    k = (0.5 * (hi - lo)[:, None] * x[None, :] + 0.5 * (hi + lo)[:, None]).ravel()
    wk = (0.5 * (hi - lo)[:, None] * w[None, :]).ravel()
    f = 2 / np.pi * k ** 2 * Pzeta(k, AS, NS, KP) * spherical_jn(L, k * a) * spherical_jn(L, k * b)
    C = 2 * np.pi ** 2 * AS * KP ** (1 - NS)  # P_zeta = C k^(ns-4)
    tail = C / (np.pi * a * b) * K ** (NS - 3) / (3 - NS) if a == b else 0.
    return np.sum(wk * f) + tail


@pytest.mark.parametrize("L", [1, 2, 5, 10])
def test_F_closed_form_vs_quadrature(L):
    r = np.array([1.0, 0.7, 0.95])
    F = F_integral_n0(r, 1, 10, NS, AS, KP)
    for i, j in ((0, 0), (0, 1), (0, 2), (2, 2), (1, 2)):
        q = _quad_F(L, r[i], r[j])
        assert abs(F[L, i, j] - q) < 1e-9 * abs(q), (L, i, j, F[L, i, j], q)
        assert F[L, i, j] == F[L, j, i]


def test_F_scaling_to_cosmological_radii():
    r = np.array([1.0, 0.7, 0.95, 1.03])
    lam = 14000.
    F1 = F_integral_n0(r, 1, 40, NS, AS, KP)
    F2 = F_integral_n0(lam * r, 1, 40, NS, AS, KP)
    assert np.allclose(F2[1:], lam ** (1 - NS) * F1[1:], rtol=1e-12, atol=0)
    assert np.all(F1[0] == 0)


def test_p_and_q_integrals_reduce_to_F():
    lmin, lmax = 2, 8
    k = np.concatenate([np.geomspace(1e-7, 1e-2, 4000)[:-1], np.arange(1e-2, 2000., 2e-3)])
    r0, rr = 1.0, np.array([0.7, 0.95])
    T = np.zeros((1, lmax + 1, len(k)))
    for l in range(lmin, lmax + 1):
        T[0, l] = spherical_jn(l, k * r0)
    Pz = Pzeta(k, AS, NS, KP)
    p = p_integral_n0(k, Pz, T, rr, lmin, lmax, lmax + 1, 1)
    q = q_integral(k, T * Pz[None, None, :], rr, lmin, lmax, 1)
    F = F_integral_n0(np.array([r0, 0.7, 0.95]), 1, lmax, NS, AS, KP)
    for l in range(lmin, lmax + 1):
        want = (-1) ** l * F[l, 0, 1:]
        assert np.all(np.abs(p[l, 0] - want) < 3e-7 * np.abs(want)), (l, p[l, 0], want)
        assert np.all(np.abs(q[l, 0] - want) < 3e-7 * np.abs(want)), (l, q[l, 0], want)
    assert np.all(p[:lmin] == 0) and np.all(q[:lmin] == 0)


def test_hyp2f1_large_L_near_one_matches_mpmath():
    """The F_L closed form needs 2F1(L+(ns-1)/2, (ns-2)/2; L+3/2; z) up to L = 3 Nside - 1 with z =
    (r_</r_>)^2 up to 1 (neighbouring radial nodes and the diagonal).  scipy's hyp2f1 returns inf /
    nan there from L ~ 170 on (0 at z = 1, L = 170); hyp2f1_F must match 30-digit mpmath."""
    import mpmath
    from polyspec_amd.k_integrals import hyp2f1_F
    zs = np.array([1.0, 1 - 1e-6, 1 - 1e-4, 1 - 8e-4, 1 - 2e-3, 1 - 3e-3, 0.99, 0.95, 0.85, 0.5])
    with mpmath.workdps(30):
        for L in (1, 50, 170, 171, 500, 1535, 3000):
            a, b, c = L + (NS - 1) / 2, (NS - 2) / 2, L + 1.5
            got = hyp2f1_F(a, b, c, zs)
            want = np.array([float(mpmath.hyp2f1(mpmath.mpf(L) + (mpmath.mpf(NS) - 1) / 2, (mpmath.mpf(NS) - 2) / 2,
                                                 mpmath.mpf(L) + mpmath.mpf(1.5), mpmath.mpf(z))) for z in zs])
            assert np.all(np.abs(got / want - 1) < 1e-8), (L, got, want)


def test_F_table_finite_at_the_nside512_workload():
    """40 Gauss-Legendre nodes on [13400, 14600] Mpc, L up to 1535 (Nside 512): every entry finite,
    symmetric, and on the diagonal equal to the Gauss-sum value of the closed form."""
    from scipy.special import gammaln, roots_legendre
    x, _ = roots_legendre(40)
    r = 600. * x + 14000.
    F = F_integral_n0(r, 1, 1535, NS, AS, KP)
    assert np.all(np.isfinite(F)) and np.allclose(F, F.transpose(0, 2, 1), rtol=1e-13, atol=0)
    L = np.arange(1, 1536)
    a, b, c = L + (NS - 1) / 2, (NS - 2) / 2, L + 1.5
    pref = 2. ** (NS - 3) * 2 * np.pi ** 2 * AS / KP ** (NS - 1)
    gauss = np.exp(gammaln(c) + gammaln(c - a - b) - gammaln(c - a) - gammaln(c - b))
    diag = pref * np.exp(gammaln(a) - gammaln(1 - b) - gammaln(c)) * gauss * r[7] ** (1 - NS)
    assert np.allclose(F[1:, 7, 7], diag, rtol=1e-11, atol=0)
