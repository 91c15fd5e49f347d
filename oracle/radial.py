"""FP64 numpy oracle of the radial-node optimizer for gNL-loc (TEST INFRASTRUCTURE ONLY).

Restates
  fisher_deriv_gNL_loc          polyspec/cython/ideal_fisher.pyx:750-787 (+ _zeta_sum_sym 811-831)
  its inputs (Gauss-Legendre mu nodes, P_l(mu) table, beam-weighted C^-1)
                                polyspec/tspec_template.py:3858-3870
  the starting radial grid      polyspec/tspec_template.py:3139-3142
  the greedy Schur-complement selection of radial nodes and their weights (gNL branch,
  no numerical weight optimisation)
                                polyspec/tspec_template.py:3532-3645, 3219-3225, 3649-3668

fisher_deriv_gnl_loc: with z^{XY}_l(r, r') = (2l + 1) sum_ij X_l^i(r) C^-1_{l,ij} Y_l^j(r')
(X, Y in {p, q}) and S^{XY}(mu; r, r') = sum_l P_l(mu) z^{XY}_l(r, r'),
  D(r, r') = pref w_r w_r' sum_mu w_mu [S^PP^3 S^QQ + 3 S^PQ S^QP S^PP^2](mu; r, r'),
  S^QP(r, r') = S^PQ(r', r),  pref = 6 / (4 pi)^2 (9/25)^2 / 2.
Pinned in tests/test_radial.py: P_l(mu) against 40-digit mpmath values at l up to 1024, the mu
quadrature against the closed form
int P_l^4 = 2 sum_L (2L+1) (l l L; 0 0 0)^4 (a single l), symmetry, and the
Schur-complement score against the direct quadratic form; sum(D) against the
Monte-Carlo gNL-loc Fisher of the estimator in the GPU tests.
"""
import numpy as np
from scipy.special import roots_legendre

PREF_GNL_LOC = 6. / (4. * np.pi) ** 2 * (9. / 25.) ** 2 / 2.


def radial_grid(r_star, r_hor, reduce_r=1):
    """Starting grid and trapezoid weights r^2 dr (tspec_template.py:3139-3142)."""
    r_raw = np.asarray(list(np.arange(1, r_star * 0.95, 50 * reduce_r)) +
                       list(np.arange(r_star * 0.95, r_hor * 1.05, 5 * reduce_r)) +
                       list(np.arange(r_hor * 1.05, r_hor + 5000, 50 * reduce_r)))
    r_init = 0.5 * (r_raw[1:] + r_raw[:-1])
    return r_init, r_init ** 2 * np.diff(r_raw)


def mu_quadrature(lmin, lmax):
    """Gauss-Legendre nodes / weights with 2 lmax + 1 points and P_l(mu) for l = lmin..lmax
    (tspec_template.py:3863-3865; lpmn's m = 0 column is the Legendre polynomial, evaluated
    like lpmn by the upward recursion (l+1) P_{l+1} = (2l+1) mu P_l - l P_{l-1})."""
    mus, w_mus = roots_legendre(2 * lmax + 1)
    legs = np.empty((len(mus), lmax - lmin + 1))
    pm1, p0 = np.zeros_like(mus), np.ones_like(mus)
    for l in range(lmax + 1):
        if l >= lmin:
            legs[:, l - lmin] = p0
        pm1, p0 = p0, ((2 * l + 1) * mus * p0 - l * pm1) / (l + 1)
    return mus, w_mus, legs


def zeta(X, Y, Cinv, lmin, lmax):
    """z_l(r, r') = (2l+1) sum_ij X[l, i, r] Cinv[i, j, l] Y[l, j, r'] for l = lmin..lmax."""
    ls = np.arange(lmin, lmax + 1)
    Xl, Yl, Cl = X[lmin:lmax + 1], Y[lmin:lmax + 1], Cinv[:, :, lmin:lmax + 1]
    return (2. * ls + 1.)[:, None, None] * np.einsum('lir,ijl,ljs->lrs', Xl, Cl, Yl)


def fisher_deriv_gnl_loc(plXs, qlXs, weights, inv_Cl_mat, legs, w_mus, lmin, lmax):
    """ideal_fisher.pyx:750-787.  plXs, qlXs: (l, pol, r) indexed by l; inv_Cl_mat: (pol, pol, l)
    (beam-weighted); legs: (mu, l - lmin); returns D (r, r')."""
    zPP = zeta(plXs, plXs, inv_Cl_mat, lmin, lmax)
    zPQ = zeta(plXs, qlXs, inv_Cl_mat, lmin, lmax)
    zQQ = zeta(qlXs, qlXs, inv_Cl_mat, lmin, lmax)
    AA = np.einsum('ml,lrs->mrs', legs, zPP)
    AB = np.einsum('ml,lrs->mrs', legs, zPQ)
    BA = np.transpose(AB, (0, 2, 1))
    BB = np.einsum('ml,lrs->mrs', legs, zQQ)
    musum = np.einsum('m,mrs->rs', w_mus, AA ** 3 * BB + 3. * AB * BA * AA ** 2)
    w = np.asarray(weights)
    return PREF_GNL_LOC * w[:, None] * w[None, :] * musum


def schur(D, inds):
    """G = D_NN - D_SN^T D_SS^-1 D_SN and w_quad = 1 + sum_cols D_SS^-1 D_SN
    (tspec_template.py:3569-3573, 3596-3600) for selected indices S = inds, N the rest."""
    n = len(D)
    notinds = [i for i in range(n) if i not in inds]
    inv = np.linalg.inv(D[np.ix_(inds, inds)])
    DSN = D[np.ix_(inds, notinds)]
    G = D[np.ix_(notinds, notinds)] - DSN.T @ inv @ DSN
    w_quad = 1. + np.sum(inv @ DSN, axis=1)
    return G, w_quad, notinds


def optimize_gnl(D, tolerance=1e-3, repeat_trials=True, stalled_iterations=np.inf, r_guess_inds=()):
    """Greedy node selection for a contact template (tspec_template.py:3532-3645, gNL branch):
    returns (inds, w_opt, score, init_score, fish) with w_opt the quadratic weights (to be
    multiplied by the starting quadrature weights)."""
    n = len(D)
    inds_init = np.arange(n)
    init_score = float(np.sum(D))
    inds = list(r_guess_inds)
    G = None
    if len(inds):
        G, w_quad, notinds = schur(D, inds)
        w_vals = w_quad
        score = float(np.sum(G))
    else:
        score, w_quad, w_vals = init_score, np.zeros(0), np.zeros(0)
    stalled_score = score
    if score / init_score >= tolerance:
        if len(inds) == 0:
            next_inds = np.argsort(np.sum(D, axis=1) ** 2 / np.diag(D))[::-1]
        else:
            next_inds = inds_init[notinds][np.argsort(np.sum(G, axis=1) ** 2 / np.diag(G))[::-1]]
        inds.append(int(next_inds[0]))
        w_vals_old, w_quad_old, score_old = w_vals, w_quad, score
        for iit, iteration in enumerate(range(len(inds), n)):
            for trial in range(len(next_inds)):
                inds[-1] = int(next_inds[trial])
                G, w_quad, notinds = schur(D, inds)
                w_vals = np.append(w_vals_old, 0.) + w_quad - np.append(w_quad_old, 0.)
                score = float(np.sum(G))
                if score < score_old or not repeat_trials:
                    break
            if score < 0:
                break
            if score / init_score < tolerance:
                break
            if iit % stalled_iterations == 0 and iit > 0:
                if abs(score - stalled_score) / stalled_score < 0.05:
                    break
                stalled_score = score
            w_vals_old, w_quad_old, score_old = w_vals, w_quad, score
            next_inds = inds_init[notinds][np.argsort(np.sum(G, axis=1) ** 2 / np.diag(G))[::-1]]
            inds.append(int(next_inds[0]))
    if G is None or len(G) == 0:
        raise RuntimeError("radial optimisation selected every node without meeting the tolerance")
    w_opt = np.asarray(w_vals, dtype=np.float64)
    fish = float(np.sum(np.outer(w_opt, w_opt) * D[np.ix_(inds, inds)]))
    return np.asarray(inds), w_opt, score, init_score, fish
