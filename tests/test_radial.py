"""Pins of the radial-node optimizer oracle (CPU) and GPU parity of the gNL-loc Fisher
derivative kernel.

  * the mu quadrature and the zeta-sum polynomial against the closed form
    int P_l^4 = 2 sum_L (2L+1) (l l L; 0 0 0)^4 for a single l (sympy's Wigner 3j);
  * D is symmetric in (r, r');
  * the Schur-complement score sum(G) equals the direct quadratic form
    (1 - v)^T D (1 - v) of the quadratic weights v, and the greedy selection meets its
    tolerance on a Fisher-derivative matrix of the synthetic transfer;
  * GPU: ps_fisher_deriv_gnl_loc against the oracle (1e-10), and sum(D) against the
    Monte-Carlo gNL-loc Fisher of the estimator (the ideal Fisher it is meant to equal).
"""
import numpy as np
import pytest

from oracle.radial import PREF_GNL_LOC, fisher_deriv_gnl_loc, mu_quadrature, optimize_gnl, radial_grid, schur


def _int_p4(l):
    from sympy.physics.wigner import wigner_3j
    return 2. * sum((2 * L + 1) * float(wigner_3j(l, l, L, 0, 0, 0)) ** 4 for L in range(0, 2 * l + 1))


@pytest.mark.parametrize("l", [2, 5, 11])
def test_mu_quadrature_single_l_closed_form(l):
    rng = np.random.default_rng(l)
    nr = 3
    P = np.zeros((l + 1, 1, nr))
    Q = np.zeros((l + 1, 1, nr))
    P[l, 0] = rng.uniform(0.5, 2, nr)
    Q[l, 0] = rng.uniform(-1, 1, nr)
    C = np.zeros((1, 1, l + 1))
    C[0, 0, l] = 0.7
    w = rng.uniform(1, 3, nr)
    _, w_mus, legs = mu_quadrature(l, l)
    D = fisher_deriv_gnl_loc(P, Q, w, C, legs, w_mus, l, l)
    t = (2 * l + 1) * 0.7
    zPP = t * np.outer(P[l, 0], P[l, 0])
    zPQ = t * np.outer(P[l, 0], Q[l, 0])
    zQQ = t * np.outer(Q[l, 0], Q[l, 0])
    want = PREF_GNL_LOC * np.outer(w, w) * (zPP ** 3 * zQQ + 3 * zPQ * zPQ.T * zPP ** 2) * _int_p4(l)
    assert np.allclose(D, want, rtol=1e-12, atol=0)


def _synthetic_tables(lmin, lmax, nr, npol, seed):
    rng = np.random.default_rng(seed)
    ls = np.arange(lmax + 1)
    r = np.linspace(13000., 15000., nr)
    # smooth in r, decaying in l: an oscillating radial kernel like j_l(k r) integrals
    P = np.zeros((lmax + 1, npol, nr))
    Q = np.zeros((lmax + 1, npol, nr))
    for i in range(npol):
        a, b = rng.uniform(0.5, 1.5, 2)
        P[lmin:, i] = (np.cos(np.outer(ls[lmin:], r) / 9000. * a) / ls[lmin:, None] ** 1.5)
        Q[lmin:, i] = (np.sin(np.outer(ls[lmin:], r) / 7000. * b) / ls[lmin:, None] ** 0.5)
    C = np.zeros((npol, npol, lmax + 1))
    for l in range(lmin, lmax + 1):
        M = rng.standard_normal((npol, npol))
        C[:, :, l] = (M @ M.T + npol * np.eye(npol)) * l * (l + 1.)
    w = r ** 2 * 5.
    return P, Q, C, w


def test_deriv_symmetric():
    lmin, lmax = 2, 20
    P, Q, C, w = _synthetic_tables(lmin, lmax, 9, 3, 1)
    _, w_mus, legs = mu_quadrature(lmin, lmax)
    D = fisher_deriv_gnl_loc(P, Q, w, C, legs, w_mus, lmin, lmax)
    assert np.allclose(D, D.T, rtol=1e-12, atol=1e-14 * np.abs(D).max())


def test_schur_score_is_the_quadratic_form():
    lmin, lmax = 2, 16
    P, Q, C, w = _synthetic_tables(lmin, lmax, 12, 1, 2)
    _, w_mus, legs = mu_quadrature(lmin, lmax)
    D = fisher_deriv_gnl_loc(P, Q, w, C, legs, w_mus, lmin, lmax)
    for inds in ([3], [0, 7], [2, 5, 9, 11]):
        G, w_quad, notinds = schur(D, inds)
        v = np.zeros(len(D))
        v[inds] = w_quad
        direct = (1 - v) @ D @ (1 - v)
        assert abs(np.sum(G) - direct) <= 1e-9 * np.sum(D)
        # and it is the minimum over weights on the selected nodes
        for _ in range(3):
            v2 = v.copy()
            v2[inds] += 1e-3 * np.random.default_rng(len(inds)).standard_normal(len(inds))
            assert (1 - v2) @ D @ (1 - v2) >= direct - 1e-12 * abs(direct)


def test_optimizer_meets_tolerance():
    lmin, lmax = 2, 24
    P, Q, C, w = _synthetic_tables(lmin, lmax, 40, 1, 3)
    _, w_mus, legs = mu_quadrature(lmin, lmax)
    D = fisher_deriv_gnl_loc(P, Q, w, C, legs, w_mus, lmin, lmax)
    inds, w_opt, score, init, fish = optimize_gnl(D, tolerance=1e-4)
    assert score / init < 1e-4
    assert len(set(inds.tolist())) == len(inds) < len(D)
    v = np.zeros(len(D))
    v[inds] = w_opt
    assert abs((1 - v) @ D @ (1 - v) - score) <= 1e-9 * init
    assert abs(fish - v @ D @ v) <= 1e-9 * init


def test_legendre_table_against_mpmath():
    import mpmath
    mpmath.mp.dps = 40
    mus, _, legs = mu_quadrature(2, 1024)
    for l in (2, 17, 500, 1024):
        for i in (0, 5, 700, 1024, 2048):
            want = float(mpmath.legendre(l, mpmath.mpf(mus[i])))
            assert abs(legs[i, l - 2] - want) < 1e-11  # rounding of the recursion grows toward |mu| = 1


def test_radial_grid_shape():
    r, w = radial_grid(13882.607764400758, 14163.032903769747)
    assert len(r) == len(w) and np.all(np.diff(r) > 0) and np.all(w > 0)
    assert 600 < len(r) < 800


def gpu_fisher_deriv(P, Q, C, w, legs, w_mus, lmin, lmax, dev):
    """ps_fisher_deriv_gnl_loc on the same tables as fisher_deriv_gnl_loc."""
    import torch
    from polyspec_amd._abi import check, load, ptr, stream_handle
    nl = lmax - lmin + 1
    npol, nr = P.shape[1], P.shape[2]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    Pd, Qd, Cd = t(P[lmin:lmax + 1]), t(Q[lmin:lmax + 1]), t(C[:, :, lmin:lmax + 1])
    Ld, Wd, wd = t(legs), t(w_mus), t(w)
    nb = int(load().ps_fisher_deriv_work_bytes(nl, npol, nr, len(w_mus)))
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    out = torch.empty((nr, nr), dtype=torch.float64, device=dev)
    check(load().ps_fisher_deriv_gnl_loc(ptr(Pd), ptr(Qd), nl, npol, nr, ptr(Cd), ptr(Ld), ptr(Wd), len(w_mus),
                                         ptr(wd), lmin, ptr(ws), nb, ptr(out), stream_handle()))
    return out.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("lmin,lmax,nr,npol", [(2, 40, 37, 3), (2, 20, 5, 1), (2, 100, 70, 2), (3, 64, 16, 3)])
def test_gpu_fisher_deriv_matches_oracle(gpu, lmin, lmax, nr, npol):
    print(f"fisher deriv lmin={lmin} lmax={lmax} nr={nr} npol={npol}", flush=True)
    P, Q, C, w = _synthetic_tables(lmin, lmax, nr, npol, nr + lmax)
    _, w_mus, legs = mu_quadrature(lmin, lmax)
    want = fisher_deriv_gnl_loc(P, Q, w, C, legs, w_mus, lmin, lmax)
    got = gpu_fisher_deriv(P, Q, C, w, legs, w_mus, lmin, lmax, gpu)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    assert np.array_equal(got, got.T)


def _template(nside, lmax, pol, nr, precision=64):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    from estimator_setup import radial_nodes, spectra, transfer
    raw, tot, beam = spectra(nside, pol)
    base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=precision)
    mask = np.ones(12 * nside ** 2)
    W = Weightings(base, mask, dict(raw))
    k, Tl = transfer(lmax, pol)
    r, w = radial_nodes(nr)
    return TSpecTemplate(base, mask, W.applySinv_ideal, ['gNL-loc'], k, Tl, 2, lmax, r_values=r,
                         r_weights={'gNL-loc': w})


@pytest.mark.gpu
@pytest.mark.parametrize("pol", [False, True])
def test_gpu_ideal_fisher_matches_monte_carlo_fisher(gpu, pol):
    """sum(D) is the ideal (full-sky, ideal S^-1) gNL-loc Fisher: it must agree with the
    estimator's Monte-Carlo Fisher within its sampling error."""
    print(f"ideal vs MC fisher pol={pol}", flush=True)
    tm = _template(16, 30, pol, 4)
    D = tm._fisher_deriv_gnl_loc(tm.r_weights['gNL-loc'])
    vals = np.array([tm.compute_fisher_contribution(seed)[0, 0] for seed in range(60)])
    mean, err = vals.mean(), vals.std(ddof=1) / np.sqrt(len(vals))
    print("ideal %.6e  MC %.6e +- %.2e" % (np.sum(D), mean, err), flush=True)
    assert abs(np.sum(D) - mean) < 5 * err + 1e-3 * abs(mean)


@pytest.mark.gpu
def test_gpu_radial_optimizer_matches_oracle(gpu):
    """optimize_radial_sampling_1d: the device Fisher derivative on the starting grid equals
    the oracle's, and the selected nodes / weights are the oracle's greedy selection on it."""
    from oracle.radial import radial_grid
    print("radial optimizer lmax=30 pol=True", flush=True)
    tm = _template(16, 30, True, 4)
    r_opt, w_opt = tm.optimize_radial_sampling_1d(tolerance=1e-3)
    r_init, qw = radial_grid(tm.r_star, tm.r_hor)
    # oracle D on the starting grid's tables (recomputed: the template now holds the optimized nodes)
    tm2 = _template(16, 30, True, 4)
    tm2.r_arr, tm2.N_r = r_init, len(r_init)
    tm2.r_weights = {'gNL-loc': qw}
    tm2._prepare_templates()
    from scipy.special import eval_legendre, roots_legendre
    mus, w_mus = roots_legendre(2 * tm2.lmax + 1)
    legs = eval_legendre(np.arange(tm2.lmin, tm2.lmax + 1)[None, :], mus[:, None])
    beam = np.asarray(tm2.base.beam)
    C = beam[:, None, :] * beam[None, :, :] * tm2.base.inv_Cl_tot_mat
    want = fisher_deriv_gnl_loc(tm2.plLXs[0], tm2.qlXs, qw, C, legs, w_mus, tm2.lmin, tm2.lmax)
    got = tm2._fisher_deriv_gnl_loc(qw)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    inds, wq, score, init, fish = optimize_gnl(want, tolerance=1e-3)
    assert np.allclose(r_opt, r_init[inds], rtol=0, atol=0)
    assert np.allclose(w_opt['gNL-loc'], wq * qw[inds], rtol=1e-6, atol=0)
    assert score / init < 1e-3
    assert abs(tm.ideal_fisher['gNL-loc'] - init) < 1e-10 * init
