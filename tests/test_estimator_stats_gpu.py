"""Statistical pins of the estimator normalisation: what the theory of the unwindowed quartic
estimator fixes independently of how its prefactors were read off the reference.

For Gaussian data analysed with the ideal S^-1 on a unit mask the estimator is optimal, so its
numerator N (t4 + t2 + t0, tspec_template.py:1338-1954) has
  (1) E[N] = 0 -- the Monte-Carlo terms t2, t0 remove the disconnected part of t4 exactly; this
      pins the ratios of the t4 / t2 / t0 prefactors (gNL-loc: 9/100*4, -54/100, +54/100 at
      tspec_template.py:1391, 1575, 1579; tauNL-loc: 1/24*12, -1/24*6/2, 1/24*3 at 1431, 1678,
      1687), since E[t4] is many standard deviations away from zero;
  (2) Cov(N) = F, the Fisher matrix -- this pins the overall normalisation of the numerator
      against the Fisher contribution (compute_fisher_contribution, 2590-3052), and for gNL-loc
      against the analytic ideal Fisher sum(D) (fisher_deriv_gNL_loc, ideal_fisher.pyx:750-787),
      which tests/test_radial.py pins against the Monte-Carlo Fisher.
Data and simulations are band-limited Gaussian maps at l <= lmax through the same map-space
S^-1, so E[N] = 0 holds exactly.  With the simulations fixed, the mean over data carries their
Monte-Carlo noise as well (about 0.04 sigma at 100 pairs, measured by swapping simulation sets
under fixed data), which the 5-standard-error mean bound covers.  Var(N) exceeds F by the Monte-Carlo noise of the disconnected
terms, O(1/N_sims), and the map-space quadrature: a 10 % allowance on top of 5 standard errors,
far below the factors of 2 or more that a misread prefactor produces.
"""
import numpy as np
import pytest
import torch

from estimator_setup import grf_alm, radial_nodes, spectra, transfer

pytestmark = pytest.mark.gpu


def _setup(nside, lmax, nr, templates):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    raw, tot, beam = spectra(nside, False)
    base = PolySpec(nside, dict(tot), beam=beam, pol=False, precision=64)
    mask = np.ones(12 * nside ** 2)
    W = Weightings(base, mask, dict(raw))
    k, Tl = transfer(lmax, False)
    r, w = radial_nodes(nr)
    tm = TSpecTemplate(base, mask, W.applySinv_ideal, templates, k, Tl, 2, lmax, r_values=r,
                       r_weights={t: w for t in templates})
    return base, tm, tot


def _maps(base, rng, tot, lmax, n):
    """n band-limited (l <= lmax) Gaussian T maps with the total spectrum, (n, Npix) on the device."""
    alm = np.stack([grf_alm(rng, tot, lmax, False)[0] for _ in range(n)])
    return base.to_map_vec(torch.from_numpy(alm).to(base.device), lmax=lmax)


def _moments(x):
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    mu = x.mean()
    d = x - mu
    var = np.mean(d ** 2) * n / (n - 1)
    se_var = np.sqrt(max(np.mean(d ** 4) - var ** 2, 0.) / n)
    return mu, var, np.sqrt(var / n), se_var


def _check(name, N, t4, F, F_err, power=10.):
    mu, var, se_mu, se_var = _moments(N)
    print("%s: E[t4]/sigma = %.1f, mean(N)/sigma = %+.3f (+- %.3f), var(N)/F = %.4f (+- %.4f stat, +- %.4f F)"
          % (name, np.mean(t4) / np.sqrt(var), mu / np.sqrt(var), se_mu / np.sqrt(var), var / F, se_var / F,
             F_err / F * var / F), flush=True)
    assert abs(np.mean(t4)) > power * np.sqrt(var), "E[t4] too small for the mean test to pin the prefactors"
    assert abs(mu) < 5 * se_mu, (name, mu, se_mu)
    tol = 5 * np.hypot(se_var / F, F_err / F * var / F) + 0.10
    assert abs(var / F - 1) < tol, (name, var / F, tol)


def test_gnl_loc_numerator_mean_zero_and_variance_equals_ideal_fisher(gpu):
    """gNL-loc alone (collapsed Monte-Carlo sums, so 100 simulation pairs cost one pass)."""
    nside, lmax, nr = 16, 24, 4
    print(f"gNL-loc statistics nside={nside} lmax={lmax} nr={nr}", flush=True)
    base, tm, tot = _setup(nside, lmax, nr, ['gNL-loc'])
    rng = np.random.default_rng(2024)
    n_pairs, n_data = 100, 1500
    sims = _maps(base, rng, tot, lmax, 2 * n_pairs)
    tm.load_sims(lambda ii: (sims[2 * ii:2 * ii + 1], sims[2 * ii + 1:2 * ii + 2]), n_pairs, input_type='map')
    data = _maps(base, rng, tot, lmax, n_data)
    N = np.array([tm.Tl_numerator(data[i:i + 1])[0] for i in range(n_data)])
    t4 = np.array([tm.Tl_numerator(data[i:i + 1], include_disconnected_term=False)[0] for i in range(200)])
    F = float(np.sum(tm._fisher_deriv_gnl_loc(tm.r_weights['gNL-loc'])))
    _check('gNL-loc', N, t4, F, 0.)


def test_local_numerators_covariance_equals_monte_carlo_fisher(gpu):
    """gNL-loc and tauNL-loc together: mean zero, and the 2x2 covariance of the numerators
    equals the Monte-Carlo Fisher matrix of compute_fisher_contribution (diagonal within the
    tolerance above, correlation coefficient within 0.1)."""
    nside, lmax, nr = 16, 20, 3
    templates = ['gNL-loc', 'tauNL-loc']
    print(f"gNL+tauNL statistics nside={nside} lmax={lmax} nr={nr}", flush=True)
    base, tm, tot = _setup(nside, lmax, nr, templates)
    rng = np.random.default_rng(77)
    n_pairs, n_data, n_fish = 100, 1000, 100
    sims = _maps(base, rng, tot, lmax, 2 * n_pairs)
    tm.load_sims(lambda ii: (sims[2 * ii:2 * ii + 1], sims[2 * ii + 1:2 * ii + 2]), n_pairs, input_type='map')
    data = _maps(base, rng, tot, lmax, n_data)
    N = np.array([tm.Tl_numerator(data[i:i + 1]) for i in range(n_data)])
    t4 = np.array([tm.Tl_numerator(data[i:i + 1], include_disconnected_term=False) for i in range(150)])
    fish = np.array([tm.compute_fisher_contribution(seed) for seed in range(n_fish)])
    F, F_err = fish.mean(0), fish.std(0, ddof=1) / np.sqrt(n_fish)
    for i, t in enumerate(templates):
        # E[t4] of tauNL-loc is about 4 standard deviations at this size: its mean test pins the
        # t4 / t2 / t0 ratios to a few per cent (gNL-loc: better than 1 %)
        _check(t, N[:, i], t4[:, i], F[i, i], F_err[i, i], power=10. if t == 'gNL-loc' else 3.)
    C = np.cov(N.T)
    rho_N = C[0, 1] / np.sqrt(C[0, 0] * C[1, 1])
    rho_F = F[0, 1] / np.sqrt(F[0, 0] * F[1, 1])
    print("correlation: numerators %.4f, Fisher %.4f" % (rho_N, rho_F), flush=True)
    assert abs(rho_N - rho_F) < 0.1
