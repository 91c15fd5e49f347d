"""FP64 oracle of the tauNL-loc branch of the radial-node optimizer (TEST INFRASTRUCTURE ONLY).

Restates, for the local template (n1 = n3 = n = 0), on the oracle SHTs and loop nests of
OracleTSpec:
  _compute_fisher_derivatives, tauNL branch     tspec_template.py:3906-3951, 3950-3973, 4062-4068
  _compute_ideal_Q4                              4071-4099
  _get_Q4_perms / _Q4_contributions (n1 = 0)     2521-2554, 2341-2436 (the p-side derivative with
                                                 r_weights / sqrt(4 pi), 2436, as written there)
  _F_PQ / _compute_PQ_vecs / _convolve_F, inds   2287-2306, 1967-1980, 2142-2162;
                                                 convolve_F_n000 tauNL_utils.pyx:966-992
  integrate_pq / integrate_pq_deriv              ideal_fisher.pyx:361-375, 395-407
  apply_ideal_weight, outer_product(_ideal)      ideal_fisher.pyx:499-522, 423-497
  _tau_direc_score                               tspec_template.py:3231-3346, 3485-3486
  the greedy selection with a Monte-Carlo score  3532-3645 (default path: no weight optimisation)

Pinned in tests/test_radial_tau_gpu.py, besides the GPU comparison: sum(D) = the score (the
per-node terms add up to Q4 when every weight is the quadrature weight), D positive
semi-definite, a zero score for the full node set in any order, and -- with a unit beam on the
unit mask, where the ideal weighting m_weight C^-1 and the estimator's Fisher weighting coincide --
the score equals the Monte-Carlo Fisher contribution of the same GRF pair.
"""
import numpy as np

from . import convolve_F_n000, integrate_pq, outer_product

FOURPI15 = (4. * np.pi) ** 1.5
OUTER_C = np.array([[1., 0., 0., -3.], [0., 1., -3., 0.], [0., -3., 9., 0.], [-3., 0., 0., 9.]])
PERMS = [([0], [0], [12.]), ([1], [1], [12.]), ([1, 0], [0, 2], [4., 4.]), ([0, 1], [1, 2], [4., 4.])]


def state(ot, a_maps):
    """P, Q maps (N_r, Npix) of both GRFs and the [P Q] vectors (nLM, N_r) of the pairings
    (0,0), (1,1), (0,1)+(1,0); a_maps harmonic at base lmax, beam deconvolved."""
    U = [np.ascontiguousarray(ot.applyAinv(a, lmax=ot.base_lmax)[:, ot.lfilt]) for a in a_maps]
    P = [ot.weighted_maps(u, ot.plLX0, 1. / np.sqrt(4. * np.pi)) for u in U]
    Q = [ot.weighted_maps(u, ot.qlXs, 1.) for u in U]
    prods = [Q[0] * P[0], Q[1] * P[1], Q[1] * P[0] + Q[0] * P[1]]
    PQ = [np.ascontiguousarray(ot._to_lm_vec(p, ot.nLmax).T[ot.nLminfilt, :]) for p in prods]
    return {'P': P, 'Q': Q, 'PQ': PQ}


def q4(ot, st, inds, w, del_w=None):
    """Q4 (4, nslot, nl) on the nodes inds with weights w (FOURPI15 included); with del_w also
    dQ4 (4, n, nslot, nl)."""
    inds = np.asarray(inds)
    n = len(inds)
    w = np.asarray(w, dtype=np.float64)
    FLL = np.ascontiguousarray(ot.FLL[:, inds][:, :, inds])
    F = []
    for k in range(3):
        conv = convolve_F_n000(np.ascontiguousarray(st['PQ'][k][:, inds]), FLL, w, ot.Ls, ot.L_indices,
                               len(ot.nLminfilt))
        F.append(ot._to_map_vec(conv.T, ot.nLmax))
    fq = np.ascontiguousarray(ot.qlXs[:, :, inds])
    fp = np.ascontiguousarray(ot.plLX0[:, :, inds])
    nslot, nl = ot.qlXs.shape[1], len(ot.ls)
    Q4 = np.zeros((4, nslot, nl), dtype=np.complex128)
    dQ = np.zeros((4, n, nslot, nl), dtype=np.complex128) if del_w is not None else None
    for q, (pq, fi, wt) in enumerate(PERMS):
        prodP = sum(wt[k] * st['P'][pq[k]][inds] * F[fi[k]] for k in range(len(pq)))
        PF = np.ascontiguousarray(ot._to_lm_vec(prodP, ot.lmax).T[ot.lminfilt])
        out = np.zeros((nslot, nl), dtype=np.complex128)
        integrate_pq(fq, PF, ot.ls, w, out)
        prodQ = sum(wt[k] * st['Q'][pq[k]][inds] * F[fi[k]] for k in range(len(pq)))
        QF = np.ascontiguousarray(ot._to_lm_vec(prodQ, ot.lmax).T[ot.lminfilt])
        integrate_pq(fp, QF, ot.ls, w / np.sqrt(4. * np.pi), out)
        Q4[q] = FOURPI15 * out
        if dQ is not None:
            # integrate_pq_deriv: out2[r, X, lm] = 0.5 w_r f[l, X, r] P_FPQ[lm, r]
            dd = np.asarray(del_w, dtype=np.float64)
            wr = w / np.sqrt(4. * np.pi)
            for r in range(n):
                dQ[q, r] = FOURPI15 * 0.5 * (dd[r] * fq[ot.ls, :, r].T * PF[:, r] + wr[r] * fp[ot.ls, :, r].T * QF[:, r])
    return Q4, dQ


def weight(ot, x):
    """apply_ideal_weight: m_weight C^-1 x over the last two axes (nslot, nl), no beam."""
    Ci = ot.invC[:, :, ot.ls]
    return ot.m_weight * np.einsum('abl,...bl->...al', Ci, x)


def derivatives(ot, grf_pairs, quad):
    """_compute_fisher_derivatives (tauNL-loc) over the GRF pairs: (mean score, mean D, states,
    full-grid (Q4, W Q4) per pair)."""
    N0 = ot.qlXs.shape[2]
    allin = np.arange(N0)
    score, D, states, Q4s = 0., 0., [], []
    for a_maps in grf_pairs:
        st = state(ot, a_maps)
        Q4, dQ = q4(ot, st, allin, quad, del_w=quad)
        WQ4 = weight(ot, Q4)
        score += outer_product(Q4.reshape(4, 1, -1), WQ4.reshape(4, 1, -1), sym=True)[0, 0] / len(grf_pairs)
        D = D + outer_product(dQ.reshape(4, N0, -1), weight(ot, dQ).reshape(4, N0, -1), sym=True) / len(grf_pairs)
        states.append(st)
        Q4s.append((Q4, WQ4))
    return float(score), D, states, Q4s


def score(ot, states, Q4s, quad, inds, w_vals, full_score=False, gradient=False):
    """_tau_direc_score for tauNL-loc: Fisher distance of the subset (and its own Fisher, or the
    gradient along w_vals), averaged over the realizations."""
    inds = np.asarray(inds)
    this_w = quad[inds] * np.asarray(w_vals, dtype=np.float64)
    t, f, g = 0., 0., np.zeros(len(inds))
    for st, (Q4f, WQ4f) in zip(states, Q4s):
        Q4w, dQ = q4(ot, st, inds, this_w, del_w=quad[inds] if gradient else None)
        WQ4w = weight(ot, Q4w)
        t += outer_product((Q4f - Q4w).reshape(4, 1, -1), (WQ4f - WQ4w).reshape(4, 1, -1), sym=True)[0, 0]
        f += outer_product(Q4w.reshape(4, 1, -1), WQ4w.reshape(4, 1, -1), sym=True)[0, 0]
        if gradient:
            g += 4. * outer_product(dQ.reshape(4, len(inds), -1), (WQ4w - WQ4f).reshape(4, 1, -1))[:, 0]
    n = len(states)
    if gradient:
        return t / n, g / n
    return (t / n, f / n) if full_score else t / n


def schur(D, inds):
    n = len(D)
    notinds = [i for i in range(n) if i not in inds]
    inv = np.linalg.inv(D[np.ix_(inds, inds)])
    DSN = D[np.ix_(inds, notinds)]
    return D[np.ix_(notinds, notinds)] - DSN.T @ inv @ DSN, 1. + np.sum(inv @ DSN, axis=1), notinds


def bfgs_test(ot, states, Q4s, quad, init_score, maxiter=10):
    """_test_inds with optimize_weights (tspec_template.py:3514-3524): BFGS on the score
    normalised by the previous score, started from the quadratic weights."""
    from scipy.optimize import minimize

    def test(inds, score_old, w_vals):
        scale = init_score if score_old == np.inf else score_old
        fun = lambda x, sc: tuple(v / sc for v in score(ot, states, Q4s, quad, inds, x, gradient=True))
        res = minimize(fun, w_vals, method='BFGS', jac=True, options={"xrtol": 1e-6, "maxiter": maxiter},
                       args=(scale,))
        return (scale * res.fun if res.fun < 1 else scale), res.x
    return test


def optimize(D, init_score, score_fn, tolerance=1e-3, repeat_trials=True, start_inds=(), test_fn=None):
    """Greedy selection (tspec_template.py:3532-3645) with trial scores score_fn(inds, w_vals), or
    test_fn(inds, score_old, w_vals) -> (score, w_vals) when the weights are refined;
    returns (inds, w_vals, score)."""
    if test_fn is None:
        test_fn = lambda inds, score_old, w_vals: (score_fn(inds, w_vals), w_vals)
    n = len(D)
    inds_init = np.arange(n)
    inds = list(start_inds)
    if len(inds):
        G, w_quad, notinds = schur(D, inds)
        score, w_vals = test_fn(list(inds), init_score, w_quad)
    else:
        G, score, w_quad, w_vals = None, init_score, np.zeros(0), np.zeros(0)
    if score / init_score >= tolerance:
        if len(inds) == 0:
            next_inds = np.argsort(np.sum(D, axis=1) ** 2 / np.diag(D))[::-1]
        else:
            next_inds = inds_init[notinds][np.argsort(np.sum(G, axis=1) ** 2 / np.diag(G))[::-1]]
        inds.append(int(next_inds[0]))
        w_vals_old, w_quad_old, score_old = w_vals, w_quad, score
        for _ in range(len(inds), n):
            for trial in range(len(next_inds)):
                inds[-1] = int(next_inds[trial])
                G, w_quad, notinds = schur(D, inds)
                w_vals = np.append(w_vals_old, 0.) + w_quad - np.append(w_quad_old, 0.)
                score, w_vals = test_fn(list(inds), score_old, w_vals)
                if score < score_old or not repeat_trials:
                    break
            if score < 0 or score / init_score < tolerance:
                break
            w_vals_old, w_quad_old, score_old = w_vals, w_quad, score
            next_inds = inds_init[notinds][np.argsort(np.sum(G, axis=1) ** 2 / np.diag(G))[::-1]]
            inds.append(int(next_inds[0]))
    if G is None or len(G) == 0:
        raise RuntimeError("radial optimisation selected every node without meeting the tolerance")
    return np.asarray(inds), np.asarray(w_vals), score
