"""tauNL-loc branch of optimize_radial_sampling_1d on the device.

The reference (tspec_template.py:3111-3672) scores trial node sets of tauNL-loc by a Monte-Carlo
ideal Fisher distance (_tau_direc_score 3231-3346, derivatives 3906-4069).  The device path
(polyspec_amd/tau_ideal.py) is compared with the FP64 oracle (oracle/radial_tau.py) on the same
starting-grid tables and GRFs: the score, the derivative matrix D, the selected nodes and their
weights, alone, with BFGS weight refinement, and with gNL-loc continuing from the tauNL-loc nodes.

Identities that fix the score independently of that restatement:
  - sum(D) = the score: with every weight at its quadrature value the per-node terms of Q4 add up
    to Q4 (exactly, at the derivative's own weights);
  - D is positive semi-definite (a Gram matrix of W^(1/2) dQ4 under the PSD permutation form);
  - the full node set, in any order, has a zero Fisher distance;
  - on the unit mask with a unit beam the ideal weighting m_weight C^-1 coincides with the
    estimator's Fisher weighting, so the score equals the Monte-Carlo Fisher contribution of the
    same GRF pair (compute_fisher_contribution, pinned by Cov(N) = F in test_estimator_stats_gpu);
  - at unit w_vals on a subset the gradient terms add up to 4 <Q4_S, W (Q4_S - Q4)>.
"""
import numpy as np
import pytest
import torch

from estimator_setup import radial_nodes, spectra, transfer

pytestmark = pytest.mark.gpu

R_STAR, R_HOR = 13882.607764400758, 14163.032903769747


def _template(nside, lmax, pol, templates):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    raw, tot, beam = spectra(nside, pol)
    beam = [np.ones_like(b) for b in beam]  # unit beam: the ideal weighting is the Fisher weighting
    base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=64)
    mask = np.ones(12 * nside ** 2)
    W = Weightings(base, mask, dict(raw))
    k, Tl = transfer(lmax, pol)
    r, w = radial_nodes(4)
    tm = TSpecTemplate(base, mask, W.applySinv_ideal, templates, k, Tl, 2, lmax, r_values=r,
                       r_weights={t: w for t in templates})
    return tm


def _grfs(tm, n_fish):
    out = []
    for seed in range(n_fish):
        pair = []
        for ii in range(2):
            a = tm.base.generate_data(seed=seed + int((1 + ii) * 1e9), output_type='harmonic', deconvolve_beam=True)
            pair.append(a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a))
        out.append(pair)
    return out


def _oracle(tm):
    from oracle.tspec import OracleTSpec
    sg = tm.start_grid
    return OracleTSpec(tm.base.Nside, tm.lmin, tm.lmax, tm.Lmin, tm.Lmax, dict(tm.base.Cl_tot),
                       np.asarray(tm.base.beam), tm.pol, ['tauNL-loc'], sg['plXs'], sg['qlXs'], sg['FLL'],
                       {'tauNL-loc': sg['weights']})


def _start_points(step):
    from oracle.radial import radial_grid
    r_init, _ = radial_grid(R_STAR, R_HOR)
    return r_init, r_init[np.arange(270, 600, step)]


@pytest.mark.parametrize("pol,step,tol,n_fish,bfgs", [(False, 15, 1e-4, 2, False), (True, 20, 1e-3, 1, False),
                                                      (False, 20, 1e-3, 1, True)])
def test_tau_optimizer_matches_oracle(gpu, pol, step, tol, n_fish, bfgs):
    from oracle import radial_tau as RT
    nside, lmax = 8, 12
    print(f"tauNL radial optimizer pol={pol} step={step} tol={tol} n_fish={n_fish} bfgs={bfgs}", flush=True)
    tm = _template(nside, lmax, pol, ['tauNL-loc'])
    r_init, pts = _start_points(step)
    r_opt, w_opt = tm.optimize_radial_sampling_1d(tolerance=tol, N_fish_optim=n_fish, initial_r_points=pts,
                                                  optimize_weights=bfgs)
    sg = tm.start_grid
    ot = _oracle(tm)
    quad = sg['weights']
    score, D, states, Q4s = RT.derivatives(ot, _grfs(tm, n_fish), quad)
    got_D = sg['fisher_deriv']['tauNL-loc']
    print("score %.6e (oracle %.6e); |dD|/max|D| %.2e" % (tm.ideal_fisher['tauNL-loc'], score,
                                                           np.abs(got_D - D).max() / np.abs(D).max()), flush=True)
    assert abs(tm.ideal_fisher['tauNL-loc'] - score) < 1e-9 * score
    assert np.abs(got_D - D).max() < 1e-9 * np.abs(D).max()
    assert np.array_equal(got_D, got_D.T)
    fn = lambda inds, w: RT.score(ot, states, Q4s, quad, inds, w)
    test_fn = RT.bfgs_test(ot, states, Q4s, quad, score) if bfgs else None
    inds, w, s = RT.optimize(D, score, fn, tolerance=tol, test_fn=test_fn)
    print("nodes %s, relative score %.3e (device %.3e)" % (inds, s / score, tm.optimization_score['tauNL-loc']),
          flush=True)
    assert np.array_equal(r_opt, pts[inds])
    # the weights solve D_SS x = D_SN 1 (Schur step): D's 1e-15 agreement becomes ~cond(D_SS) 1e-16
    tol_w = 1e-5 if bfgs else 1e-6
    dw = np.abs(w_opt['tauNL-loc'] - w * quad[inds]).max() / np.abs(w * quad[inds]).max()
    print("weights: max |dw| / max |w| = %.2e, cond(D_SS) = %.2e" % (dw, np.linalg.cond(D[np.ix_(inds, inds)])),
          flush=True)
    assert dw < tol_w
    assert abs(tm.optimization_score['tauNL-loc'] - s / score) < tol_w * max(s / score, 1e-3)
    if not bfgs:
        assert s / score < tol


def test_tau_score_identities(gpu):
    from polyspec_amd import kernels as K
    from polyspec_amd.tau_ideal import TauIdeal
    nside, lmax = 8, 12
    print("tauNL ideal score identities", flush=True)
    tm = _template(nside, lmax, True, ['tauNL-loc'])
    # the ideal score on the template's own nodes and weights, against its Fisher contribution
    tm.quad_weights_1d = tm.r_weights['tauNL-loc'].copy()
    ti = TauIdeal(tm)
    a_maps = _grfs(tm, 1)[0]
    st = ti.add_realization(a_maps)
    score, D = ti.derivative(st)
    lf = tm.base.l_arr <= tm.lmax
    fish = tm.fisher_contribution_from_grfs([np.ascontiguousarray(a[:, lf]) for a in a_maps])[0, 0]
    print("ideal score %.8e, Fisher contribution %.8e, sum D %.8e" % (score, fish, D.sum()), flush=True)
    assert abs(score - fish) < 1e-9 * abs(fish)
    assert abs(D.sum() - score) < 1e-10 * score
    assert np.array_equal(D, D.T)
    assert np.linalg.eigvalsh(D).min() > -1e-10 * np.abs(D).max()
    n = tm.N_r
    assert ti.score(np.arange(n)[::-1], np.ones(n)) < 1e-12 * score
    # gradient terms at unit w_vals on a subset
    S = [2, 0]
    _, g = ti.score(S, np.ones(len(S)), gradient=True)
    Q4w, _ = ti.q4(st, S, ti.quad[S])
    WQ4w = ti.weight(Q4w)
    Q4f, WQ4f = ti.Q4[0]
    want = 4. * float(K.outer_product(Q4w.view(4, 1, -1), (WQ4w - WQ4f).view(4, 1, -1))[0, 0])
    assert abs(g.sum() - want) < 1e-9 * abs(want)


def test_gnl_continues_from_tau_nodes(gpu):
    """Both templates: tauNL-loc first, gNL-loc starting from its nodes; the tauNL-loc weights are
    zero on the nodes gNL-loc adds (tspec_template.py:3168-3170, 3649-3668)."""
    from oracle import radial_tau as RT
    from oracle.radial import optimize_gnl
    nside, lmax, tol = 8, 12, 1e-3
    print("tauNL + gNL radial optimizer", flush=True)
    tm = _template(nside, lmax, True, ['gNL-loc', 'tauNL-loc'])
    r_init, pts = _start_points(20)
    r_opt, w_opt = tm.optimize_radial_sampling_1d(tolerance=tol, initial_r_points=pts)
    sg = tm.start_grid
    ot = _oracle(tm)
    quad = sg['weights']
    score, D, states, Q4s = RT.derivatives(ot, _grfs(tm, 1), quad)
    inds_t, w_t, _ = RT.optimize(D, score, lambda i, w: RT.score(ot, states, Q4s, quad, i, w), tolerance=tol)
    inds_g, w_g, _, _, _ = optimize_gnl(sg['fisher_deriv']['gNL-loc'], tolerance=tol, r_guess_inds=list(inds_t))
    print("tau nodes %s, all nodes %s" % (inds_t, inds_g), flush=True)
    assert np.array_equal(inds_g[:len(inds_t)], inds_t)
    assert np.array_equal(r_opt, pts[inds_g])
    wg = w_g * quad[inds_g]
    assert np.abs(w_opt['gNL-loc'] - wg).max() < 1e-6 * np.abs(wg).max()
    wt = np.zeros(len(inds_g))
    wt[:len(inds_t)] = w_t * quad[inds_t]
    assert np.abs(w_opt['tauNL-loc'] - wt).max() < 1e-6 * np.abs(wt).max()
