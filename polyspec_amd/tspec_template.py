"""Local trispectrum templates (gNL-loc, tauNL-loc) on the MI355X backend.

Mirrors the gNL-loc / tauNL-loc branches of polyspec/tspec_template.py
(class TSpecTemplate): same constructor arguments, the same numerator
(Tl_numerator, tspec_template.py:1338-1954), Fisher contribution
(compute_fisher_contribution, 2590-3052), compute_fisher (3054-3072) and
Tl_unwindowed, with the same prefactors.  Everything after applySinv runs on
the device: radial weighting -> batched synthesis -> fused pixel reductions,
and product maps -> batched adjoint transforms -> radial / F_L sums.

Two evaluation modes for the Monte-Carlo (disconnected) terms:
  * batch_sims=True (default): the simulation pairs are processed in batched
    launches once, at generate_sims / load_sims time, and collapsed into
    U_r = sum_s P_s Q_s, V_r = sum_s P_s^2 (gNL-loc) and sum_s A_ss
    (tauNL-loc); per numerator the gNL-loc two-field term is then one fused
    pass (exact algebra of gnl_loc_disc_sum, tauNL_utils.pyx:275-285, summed
    over s).  The tauNL-loc A_sd terms remain per simulation (they are
    quadratic in the simulation) and are batched over simulations.
  * batch_sims=False: the reference's per-simulation loop
    (tspec_template.py:1542-1687) verbatim, one simulation at a time.
"""
import time

import numpy as np
import torch

from . import kernels as K
from .k_integrals import F_integral_n0, Pzeta as _Pzeta, jl_kr, p_integral_n0, q_integral

SUPPORTED = ('gNL-loc', 'tauNL-loc')
FOURPI15 = (4. * np.pi) ** 1.5


def _round_up(x, m):
    return (x + m - 1) // m * m


class TSpecTemplate:
    """Template trispectrum estimator restricted to the local templates (tspec_template.py:41-240)."""

    def __init__(self, base, mask, applySinv, templates, k_arr, Tl_arr, lmin, lmax, Lmin=None, Lmax=None,
                 Lmin_lens=None, Lmax_lens=None, ns=0.96, As=2.1e-9, k_pivot=0.05, r_values=[], r_weights={},
                 rtau_values=[], rtau_weights={}, C_phi=[], C_lens_weight={}, K_coll=0.1, k_coll=0.1, r_star=None,
                 r_hor=None, batch_sims=True, sim_batch=None, capture_graph=False, sim_memory_gb=None,
                 fuse_gnl=True):
        for t in templates:
            if t not in SUPPORTED:
                raise NotImplementedError("Template %s is not provided by the hip backend (only %s)" % (t, SUPPORTED))
        self.base = base
        self.mask = mask
        self.applySinv = applySinv
        self.pol = self.base.pol
        self.templates = list(templates)
        self.k_arr = np.asarray(k_arr, dtype=np.float64)
        self.lmin = lmin
        self.lmax = lmax
        self.ns, self.As, self.k_pivot = ns, As, k_pivot
        self.batch_sims = batch_sims
        self.sim_batch = sim_batch
        self.sim_batch_tau = None  # simulations per A_sd batch of the tauNL-loc t2 term (None: 256 // N_r)
        # capture_graph: with batch_sims, the device part of each numerator (S^-1, radial
        # filtering, syntheses, fused sums) is captured once into a HIP graph and replayed;
        # applySinv must then be capturable (no host synchronisation, e.g. applySinv_ideal)
        self.capture_graph = capture_graph
        # sim_memory_gb: device memory for the filtered simulation maps that the batched tauNL-loc
        # terms revisit on every numerator (default: 70 % of the free memory when they are made);
        # simulations beyond it keep only their S^-1-filtered harmonics and are re-synthesised
        self.sim_memory_gb = sim_memory_gb
        # fuse_gnl: when only the gNL-loc sum of the data's P and Q maps is needed (gNL-loc alone,
        # batched simulations or no disconnected term) and the FP32 plan has staged ring kernels,
        # the ring stage of the synthesis adds the sum directly and the maps are never written
        self.fuse_gnl = fuse_gnl
        self._graph = None
        self._sims_version = 0
        self.kernel_events = None  # set to a list to time the Legendre stage of every filter call

        assert self.lmax <= base.lmax, "Maximum l can't be larger than HEALPix resolution!"
        assert self.lmin >= 2, "Minimum l can't be less than 2"
        self.Lmin = 1 if Lmin is None else Lmin
        self.Lmax = min([base.lmax, 2 * self.lmax]) if Lmax is None else Lmax
        assert self.Lmin >= 1, "Minimum L can't be less than 1"
        assert self.Lmax <= 2 * self.lmax, "Lmax <= 2*lmax by the triangle conditions"
        assert self.Lmax <= base.lmax, "Maximum L can't be larger than HEALPix resolution"

        # l / L bookkeeping (tspec_template.py:115-125, 437-451) for nmax = nmax_F = 0
        self.lfilt = (self.base.l_arr >= self.lmin) & (self.base.l_arr <= self.lmax)
        self.ls, self.ms = self.base.l_arr[self.lfilt], self.base.m_arr[self.lfilt]
        self.m_weight = np.asarray(self.base.m_weight[self.lfilt], order='C')
        self.Cinv = np.asarray(self.base.inv_Cl_tot_lm_mat[:, :, self.lfilt], order='C')
        self.beam_lm = np.asarray(self.base.beam_lm)[:, self.lfilt]
        self.nmax = 0
        self.nmax_F = 0
        self.nLmax = self.Lmax
        self.nLmin = self.Lmin
        self.nLfilt = (self.base.l_arr >= self.nLmin) & (self.base.l_arr <= self.nLmax)
        self.Ls, self.Ms = self.base.l_arr[self.nLfilt], self.base.m_arr[self.nLfilt]
        self.nLminfilt = self.base.l_arr[self.base.l_arr <= self.nLmax] >= self.nLmin
        self.lminfilt = self.base.l_arr[self.base.l_arr <= self.lmax] >= self.lmin
        self.p_inds = [0]

        m = np.asarray(mask, dtype=np.float64)
        if m.ndim > 1:
            assert len(m[0]) == self.base.Npix, f'Mask has incorrect shape: {m.shape}'
        else:
            assert len(m) == self.base.Npix, f'Mask has incorrect shape: {m.shape}'
        self.ones_mask = bool(np.std(m) < 1e-12 and np.abs(np.mean(m) - 1) < 1e-12)
        self.r_star = 13882.607764400758 if r_star is None else r_star
        self.r_hor = 14163.032903769747 if r_hor is None else r_hor

        self.Tl_arr = np.asarray(Tl_arr, dtype=np.float64)
        if not self.pol:
            assert len(self.Tl_arr) == 1, "Transfer function should contain only T components"
        else:
            assert len(self.Tl_arr) == 2, "Transfer function should contain both T and E components"
        assert len(self.Tl_arr[0]) >= self.lmax + 1, "Transfer function must contain all ell modes of interest"
        assert self.Tl_arr[0][:2].sum() == 0., "Transfer function should return zero for ell = 0, 1"
        assert len(self.Tl_arr[0][0]) == len(k_arr), "Transfer function must be computed on the input k-grid"

        self.timers = {x: 0. for x in ['precomputation', 'numerator', 'fisher', 'fish_grfs', 'fish_outer',
                                       'fish_weighting', 'fish_convolve', 'fish_deriv', 'fish_products', 'Sinv', 'Ainv',
                                       'map_transforms', 'gNL_summation', 'tauNL_products', 'tauNL_summation']}
        self.base.time_sht = 0.
        if len(r_values) == 0:
            print("# No input radial sampling points supplied\n")
            self.N_r = 0
        else:
            for t in self.templates:
                assert t in r_weights.keys(), "Must supply weight for template %s" % t
            self.r_arr = np.asarray(r_values, dtype=np.float64)
            self.N_r = len(self.r_arr)
            self.r_weights = {k: np.asarray(v, dtype=np.float64) for k, v in r_weights.items()}
            self._prepare_templates()

    def Pzeta(self, k):
        """Primordial curvature power spectrum of the template (a method, not a stored closure,
        so that a template holding device buffers is freed as soon as it is dropped)."""
        return _Pzeta(k, self.As, self.ns, self.k_pivot)

    # ------------------------------------------------------------ precompute

    def _prepare_templates(self):
        """p_l^X(r), q_l^X(r) and F_L(r, r') (tspec_template.py:520-708, n = 0 terms)."""
        t0 = time.time()
        self._drop_sim_state()
        nslot = 1 + 2 * self.pol
        jl = jl_kr(np.arange(self.lmin, self.lmax + 1), self.k_arr, self.r_arr)  # shared by p and q
        self.plLXs = p_integral_n0(self.k_arr, self.Pzeta(self.k_arr), self.Tl_arr, self.r_arr, self.lmin, self.lmax,
                                   self.lmax + 1, nslot, jl=jl)[None]
        self.qlXs = q_integral(self.k_arr, self.Tl_arr, self.r_arr, self.lmin, self.lmax, nslot, jl=jl)
        del jl
        if 'tauNL-loc' in self.templates:
            self.FLLs = F_integral_n0(self.r_arr, self.Lmin, self.Lmax, self.ns, self.As, self.k_pivot)[None]
        self._upload_tables()
        self.timers['precomputation'] += time.time() - t0

    def _drop_sim_state(self):
        """Forget everything derived from the previous radial nodes: the processed simulations
        (their P/Q stacks and Monte-Carlo sums are per node), the cached t0 and a captured
        numerator graph (it holds the old tables and workspaces).  Simulations must be
        generated or loaded again after the nodes change."""
        for name in ('t0_num', 'preload', 'proc_maps', 'U_d', 'V_d', 'D_t', '_M_t', 'load_sim_data', 'N_it', '_sim_P', '_sim_Q',
                     '_sim_h'):
            if hasattr(self, name):
                delattr(self, name)
        self._sims_version += 1
        self._graph = None

    def _upload_tables(self):
        dev, rdt = self.base.device, self.base.rdtype
        self.fP_d = torch.from_numpy(np.ascontiguousarray(self.plLXs[self.nmax])).to(dev, rdt)
        self.fQ_d = torch.from_numpy(np.ascontiguousarray(self.qlXs)).to(dev, rdt)
        if hasattr(self, 'FLLs'):
            self.FLL_d = torch.from_numpy(np.ascontiguousarray(self.FLLs[self.nmax_F])).to(dev)
        self.w_d = {t: torch.from_numpy(self.r_weights[t]).to(dev) for t in self.templates}
        self.w_tau_sqrt = {t: torch.from_numpy(self.r_weights[t] / np.sqrt(4. * np.pi)).to(dev)
                           for t in self.templates}
        self.mw_d = torch.from_numpy(1. + (self.base.m_arr[self.base.l_arr <= self.lmax] > 0)).to(dev, rdt)
        self._lmaj = {}
        self._gnl_tables = None

    def _gnl_balanced_tables(self, P_weight=1. / np.sqrt(4. * np.pi)):
        """Radial tables for the fused gNL-loc sum: fP s_r, fQ / s_r and 1 / s_r^2 (device).  The
        fused ring stage takes P_r and Q_r as the two real columns of one complex FFT, whose absolute
        rounding error they share; Q_r is typically orders of magnitude larger than P_r (p_l carries
        the primordial power, q_l does not: 700x at Nside 512), which left the FP32 P maps with
        relative errors of ~1e-4.  s_r^2 = rms(Q_r) / rms(P_r) balances the two: measured on the
        simulations' maps when they are processed (_collapse_finish), else estimated here from the
        tables with the power (2l + 1) [C^-1]_XX(l) of S^-1 d for a Gaussian full sky."""
        if self._gnl_tables is None:
            npol = 2 if self.pol else 1
            fP = np.asarray(self.plLXs[self.nmax], dtype=np.float64)[:self.lmax + 1, :npol] * P_weight
            fQ = np.asarray(self.qlXs, dtype=np.float64)[:self.lmax + 1, :npol]
            l = np.arange(self.lmax + 1)
            icl = self.base.inv_Cl_tot_mat
            pw = np.stack([(2 * l + 1.) * np.abs(icl[X, X, l]) for X in range(npol)], axis=1)  # (l, X)
            pw[l < self.lmin] = 0.
            num = np.einsum('lx,lxr->r', pw, fQ ** 2)
            den = np.einsum('lx,lxr->r', pw, fP ** 2)
            ok = (num > 0) & (den > 0) & np.isfinite(num) & np.isfinite(den)
            s2 = np.where(ok, np.sqrt(np.where(ok, num, 1.) / np.where(ok, den, 1.)), 1.)
            sr = np.sqrt(s2)
            dev, rdt = self.base.device, self.base.rdtype
            self._gnl_tables = (
                torch.from_numpy(np.ascontiguousarray(self.plLXs[self.nmax] * sr)).to(dev, rdt),
                torch.from_numpy(np.ascontiguousarray(self.qlXs / sr)).to(dev, rdt),
                torch.from_numpy(1. / s2).to(dev, torch.float64))
        return self._gnl_tables

    # ------------------------------------------------- radial-node optimizer
    def _fisher_deriv_gnl_loc(self, weights, events=None):
        """Exact gNL-loc Fisher derivative D(r, r') on the device (fisher_deriv_gNL_loc,
        ideal_fisher.pyx:750-787) with the inputs of _compute_fisher_derivatives
        (tspec_template.py:3858-3870): 2 lmax + 1 Gauss-Legendre mu nodes, P_l(mu) for
        l = lmin..lmax, the beam-weighted inverse total spectrum and the weights."""
        from scipy.special import roots_legendre
        from ._abi import check, load, ptr, stream_handle
        mus, w_mus = roots_legendre(2 * self.lmax + 1)
        # P_l(mu) by the upward three-term recursion (l+1) P_{l+1} = (2l+1) mu P_l - l P_{l-1}
        legs = np.empty((len(mus), self.lmax - self.lmin + 1))
        pm1, p0 = np.zeros_like(mus), np.ones_like(mus)
        for l in range(self.lmax + 1):
            if l >= self.lmin:
                legs[:, l - self.lmin] = p0
            pm1, p0 = p0, ((2 * l + 1) * mus * p0 - l * pm1) / (l + 1)
        beam = np.asarray(self.base.beam)
        Cinv = (beam[:, None, :] * beam[None, :, :] * self.base.inv_Cl_tot_mat)[:, :, self.lmin:self.lmax + 1]
        dev = self.base.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
        P = t(self.plLXs[self.nmax][self.lmin:self.lmax + 1])
        Q = t(self.qlXs[self.lmin:self.lmax + 1])
        nl, npol, nr = P.shape
        Cd, Ld, Wd, wd = t(Cinv), t(legs), t(w_mus), t(weights)
        nb = int(load().ps_fisher_deriv_work_bytes(nl, npol, nr, len(mus)))
        ws = torch.empty(nb, dtype=torch.uint8, device=dev)
        out = torch.empty((nr, nr), dtype=torch.float64, device=dev)
        if events is not None:  # device time of the derivative's kernels, on torch's current stream
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        check(load().ps_fisher_deriv_gnl_loc(ptr(P), ptr(Q), nl, npol, nr, ptr(Cd), ptr(Ld), ptr(Wd), len(mus),
                                             ptr(wd), self.lmin, ptr(ws), nb, ptr(out), stream_handle()))
        if events is not None:
            e1.record()
            events.append((e0, e1))
        D = out.cpu().numpy()
        del ws
        return D

    def _compute_fisher_derivatives(self, templates, N_fish_optim=None, verb=False):
        """Fisher derivatives on the current (starting) grid, tspec_template.py:3842-4069 for the
        local templates.  gNL-loc exactly: (sum D, D).  tauNL-loc from N_fish_optim pairs of
        GRFs (seeds seed + (1 + ii) 1e9, harmonic, beam deconvolved): (mean score, mean D,
        TauIdeal), the last holding each realization's maps and full-grid Q4 for the score."""
        from .tau_ideal import TauIdeal
        for t in templates:
            if t not in SUPPORTED:
                raise Exception("Template %s not implemented!" % t)
        output = {}
        t0 = time.time()
        if 'gNL-loc' in templates:
            if verb:
                print("\tComputing gNL-loc Fisher matrix derivative exactly")
            D = self._fisher_deriv_gnl_loc(self.quad_weights_1d)
            output['gNL-loc'] = (np.sum(D), D)
        self.timers['fish_deriv'] += time.time() - t0
        if 'tauNL-loc' in templates:
            n_it = 1 if N_fish_optim is None else int(N_fish_optim)
            ti = TauIdeal(self)
            score, D = 0., 0.
            for seed in range(n_it):
                print("\nComputing Fisher matrix derivatives from realization %d of %d" % (seed + 1, n_it))
                t0 = time.time()
                a_maps = [self.base.generate_data(seed=seed + int((1 + ii) * 1e9), output_type='harmonic',
                                                  deconvolve_beam=True) for ii in range(2)]
                self.timers['fish_grfs'] += time.time() - t0
                t0 = time.time()
                st = ti.add_realization(a_maps)
                if verb:
                    print("Computing tauNL-loc Fisher matrix derivative")
                s1, D1 = ti.derivative(st)
                score += s1 / n_it
                D = D + D1 / n_it
                self.timers['fish_deriv'] += time.time() - t0
            output['tauNL-loc'] = (score, D, ti)
        return output

    @staticmethod
    def _schur(D, inds):
        """G = D_NN - D_SN^T D_SS^-1 D_SN and w_quad = 1 + row sums of D_SS^-1 D_SN
        (tspec_template.py:3569-3573, 3596-3600)."""
        notinds = np.setdiff1d(np.arange(len(D)), inds)
        inv = np.linalg.inv(D[np.ix_(inds, inds)])
        DSN = D[np.ix_(inds, notinds)]
        X = inv @ DSN
        return D[np.ix_(notinds, notinds)] - DSN.T @ X, 1. + np.sum(X, axis=1), notinds

    def optimize_radial_sampling_1d(self, reduce_r=1, tolerance=1e-3, N_fish_optim=1, stalled_iterations=np.inf,
                                    N_split=None, split_index=None, initial_r_points=None, verb=False,
                                    optimize_weights=False, optimize_skip=0, repeat_trials=True, r_guess=[],
                                    weight_guess={}):
        """Radial sampling points and weights by greedy optimisation of the Fisher distance
        (tspec_template.py:3111-3672).  The Fisher derivative on the starting grid is computed on
        the device (gNL-loc exactly; tauNL-loc from N_fish_optim GRF pairs), the nodes are then
        chosen by the reference's Schur-complement selection, templates in the reference's order
        (tauNL first) and each continuing from the nodes of the previous one.  A trial's score
        is sum(G) for gNL-loc and, for tauNL-loc, the Monte-Carlo Fisher distance of the trial
        subset (device work per trial); optimize_weights / optimize_skip refine the tauNL-loc
        weights by BFGS on that score and its gradient.  Returns (r_arr, r_weights) and re-runs
        the k-space precompute on the selected nodes."""
        from scipy.optimize import minimize
        t_init = time.time()
        assert reduce_r > 0, "reduce_r parameter must be positive"
        if reduce_r < 0.5:
            print("## Caution: very dense r-sampling requested; computation may be very slow")
        if reduce_r > 3:
            print("## Caution: very sparse r-sampling requested; computation may be inaccurate")
        # starting grid (tspec_template.py:3139-3142)
        r_raw = np.asarray(list(np.arange(1, self.r_star * 0.95, 50 * reduce_r)) +
                           list(np.arange(self.r_star * 0.95, self.r_hor * 1.05, 5 * reduce_r)) +
                           list(np.arange(self.r_hor * 1.05, self.r_hor + 5000, 50 * reduce_r)))
        r_init = 0.5 * (r_raw[1:] + r_raw[:-1])
        self.quad_weights_1d = r_init ** 2 * np.diff(r_raw)
        if initial_r_points is not None:
            assert split_index is None, "Cannot specify both initial_r_points and index_split"
            assert len(initial_r_points) == len(np.unique(initial_r_points)), \
                "initial_r_points cannot contain repeated points"
            sel = np.asarray([np.where(r == r_init)[0][0] for r in initial_r_points])
            r_init, self.quad_weights_1d = r_init[sel], self.quad_weights_1d[sel]
        elif N_split is not None:
            print("Partitioning sampling grid into %d pieces" % N_split)
            r_init = r_init[split_index::N_split]
            self.quad_weights_1d = self.quad_weights_1d[split_index::N_split]
        self.r_arr = r_init
        self.N_r = len(r_init)
        self.r_weights = {t: self.quad_weights_1d.copy() for t in self.templates}
        print("# Computing k integrals with fiducial radial grid")
        self._prepare_templates()

        ordered = [t for t in self.templates if 'tauNL' in t] + [t for t in self.templates if 'tauNL' not in t]
        inds = []
        for rg in r_guess:
            index = np.where(rg == r_init)[0]
            if len(index) == 0:
                raise Exception("r_guess was computed with different settings! All points in r_guess must be in r_init.")
            inds.append(int(index[0]))
        inds_init = np.arange(self.N_r)
        inds_guess = list(inds)
        if len(weight_guess.keys()) > 0:
            assert len(inds) > 0, "Must supply r_guess!"
            if verb:
                print("Starting from a set of %d radial points" % len(inds))
            for t in ordered:
                if 'tauNL' in t:
                    assert t in weight_guess.keys(), "weight_guess does not contain template %s" % t
        if verb:
            print("Computing all Fisher matrix derivatives")
        derivs = self._compute_fisher_derivatives(ordered, N_fish_optim=N_fish_optim, verb=verb)
        # the starting grid and its tables, kept for diagnostics once the nodes are replaced
        self.start_grid = {'r': r_init.copy(), 'weights': self.quad_weights_1d.copy(),
                           'plXs': self.plLXs[self.nmax].copy(), 'qlXs': self.qlXs.copy(),
                           'FLL': self.FLLs[self.nmax_F].copy() if hasattr(self, 'FLLs') else None,
                           'fisher_deriv': {t: derivs[t][1] for t in derivs}}
        if not hasattr(self, 'ideal_fisher'):
            self.ideal_fisher = {}
        for t in ordered:
            self.ideal_fisher[t] = float(derivs[t][0])
        self.optimization_score = {}
        r_weights = {}
        w_opt = np.zeros(0)
        for template in ordered:
            init_score, D = float(derivs[template][0]), derivs[template][1]
            tau = derivs[template][2] if 'tauNL' in template else None
            if verb:
                print("\nRunning optimization for template %s%s" %
                      (template, " (numerically maximizing each step)" if tau is not None and optimize_weights else ""))
                print("Initial score: %.2e" % init_score)

            def test_inds(sel, score_old, w_vals, G, iit):
                """_test_inds (3514-3536): the trial's score, with the weight optimisation."""
                if tau is None:
                    return float(np.sum(G)), w_vals
                fun = lambda x, scale: tau.score(sel, x, scale, gradient=True)
                if optimize_weights:
                    scale = init_score if score_old == np.inf else score_old
                    res = minimize(fun, w_vals, method='BFGS', jac=True, options={"xrtol": 1e-6, "maxiter": 10},
                                   args=(scale,))
                    w_vals = res.x
                    return (scale * res.fun if res.fun < 1 else scale), w_vals
                score = tau.score(sel, w_vals)
                if (optimize_skip != 0 and iit % optimize_skip == 0 and score / init_score > tolerance and iit > 0
                        and score < score_old):
                    res = minimize(fun, w_vals, method='BFGS', jac=True, options={"xrtol": 1e-6, "maxiter": 25},
                                   args=(score,))
                    w_vals = res.x
                    if res.fun < 1:
                        score = score * res.fun
                    if verb:
                        print("Optimization reduced score by %.2f" % res.fun)
                return score, w_vals

            G = None
            if len(inds):
                G, w_quad, notinds = self._schur(D, inds)
                if template in weight_guess.keys():
                    n_guess = int(np.sum(weight_guess[template] != 0))
                    _, w_quad_guess, _ = self._schur(D, inds_guess[:n_guess])
                    w_vals = w_quad + np.concatenate([weight_guess[template][:n_guess] /
                                                      self.quad_weights_1d[inds_guess[:n_guess]] - w_quad_guess,
                                                      np.zeros(len(inds) - n_guess)])
                else:
                    w_vals = w_quad
                score, w_vals = test_inds(list(inds), init_score, w_vals, G, 0)
                if verb:
                    print("Unoptimized relative score: %.2e" % (score / init_score))
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
                for iit, iteration in enumerate(range(len(inds), self.N_r)):
                    for trial in range(len(next_inds)):
                        inds[-1] = int(next_inds[trial])
                        G, w_quad, notinds = self._schur(D, inds)
                        w_vals = np.append(w_vals_old, 0.) + w_quad - np.append(w_quad_old, 0.)
                        score, w_vals = test_inds(list(inds), score_old, w_vals, G, iit)
                        if verb:
                            print("Iteration %d, trial %d, relative score: %.2e, old score: %.2e" %
                                  (iteration, trial, score / init_score, score_old / init_score))
                        if score < score_old or not repeat_trials:
                            break
                    if score < 0:
                        print("## Score is negative; this indicates a numerical error!")
                        break
                    if score / init_score < tolerance:
                        break
                    if iit % stalled_iterations == 0 and iit > 0:
                        if np.abs(score - stalled_score) / stalled_score < 0.05:
                            print("Stalled at iteration %d" % iteration)
                            break
                        stalled_score = score
                    w_vals_old, w_quad_old, score_old = w_vals, w_quad, score
                    next_inds = inds_init[notinds][np.argsort(np.sum(G, axis=1) ** 2 / np.diag(G))[::-1]]
                    inds.append(int(next_inds[0]))
            if G is None or len(G) == 0:
                raise Exception("Failed to converge after %d iterations; this indicates a bug!" % N_fish_optim)
            if verb:
                print("\nScore threshold met with %d indices" % len(inds))
            w_opt = np.asarray(w_vals, dtype=np.float64).copy()
            if tau is None:
                fish = float(np.sum(np.outer(w_opt, w_opt) * D[np.ix_(inds, inds)]))
            else:
                score, fish = tau.score(list(inds), w_opt, full_score=True)
            if verb:
                print("Ideal %s Fisher: %.4e (initial), %.4e (optimized). Relative score: %.2e\n" %
                      (template, init_score, fish, score / init_score))
            self.optimization_score[template] = score / init_score
            r_weights[template] = w_opt * self.quad_weights_1d[inds]
        del derivs
        self.r_arr = r_init[inds]
        self.N_r = len(self.r_arr)
        self.r_weights = {}
        for template in ordered:
            # weights of the nodes each template selected, zero on nodes added after it
            self.r_weights[template] = np.zeros(len(w_opt))
            self.r_weights[template][:len(r_weights[template])] = r_weights[template]
        if verb:
            print("Computing k integrals with optimized radial grid")
        self._prepare_templates()
        print("\nOptimization complete after %.2f seconds" % (time.time() - t_init))
        return self.r_arr, self.r_weights

    # ---------------------------------------------------------------- timing
    def report_timings(self):
        print("\n## Timings ##")
        for k, v in self.timers.items():
            if v > 0:
                print("%s: %.2fs" % (k, v))
        print("\nNumber of SHTs: %d forward, %d reverse; SHT time %.2fs" % (self.base.n_SHTs_forward,
                                                                           self.base.n_SHTs_reverse,
                                                                           self.base.time_sht))

    def reset_timings(self):
        for k in self.timers:
            self.timers[k] = 0.
        self.base.time_sht = 0.
        self.base.n_SHTs_forward = 0
        self.base.n_SHTs_reverse = 0

    # ------------------------------------------------------------ filtering
    def _h_dev(self, h):
        """applySinv output (nf, nlm(lmax)) -> device (npol_h, nlm(lmax)) of the backend dtype."""
        if not isinstance(h, torch.Tensor):
            h = torch.from_numpy(np.ascontiguousarray(h))
        h = h.to(self.base.device, self.base.cdtype)
        npol = 2 if self.pol else 1
        return h[:npol].contiguous()

    def _lmajor_buffer(self, ncol, lmax):
        plan = self.base.plan(lmax, 0)
        ld = _round_up(ncol, plan.lmajor_multiple)
        key = (ld, lmax)
        buf = self._lmaj.get(key)
        if buf is None:
            buf = torch.zeros((plan.nlm, ld), dtype=self.base.cdtype, device=self.base.device)
            self._lmaj[key] = buf
        return buf, plan

    def _filter(self, hs, P_weight=1. / np.sqrt(4. * np.pi)):
        """P and Q maps of each filtered field (_apply_all_filters 'p00' and 'q',
        tspec_template.py:1036-1095 -> 799-855): one batched synthesis.

        Returns (len(hs), 2, N_r, Npix): [:, 0] = P (weighted by P_weight), [:, 1] = Q."""
        t0 = time.time()
        nr = self.N_r
        ncol = 2 * nr * len(hs)
        buf, plan = self._lmajor_buffer(ncol, self.lmax)
        for i, h in enumerate(hs):
            K.weight_alm_pq_lmajor(self.fP_d, self.fQ_d, h, self.lmax, self.lmin, P_weight, 1., buf,
                                   col_offset=2 * nr * i)
        if self.kernel_events is not None:
            # split stages with device events around the Legendre stage (kernel timing)
            key = ('phase', ncol)
            ph = self._lmaj.get(key)
            if ph is None:
                ph = plan.phase_buffer(ncol)
                self._lmaj[key] = ph
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.legendre_synthesis_lmajor(buf, ncol, ph)
            e1.record()
            self.kernel_events.append((ncol, e0, e1))
            maps = plan.ring_synthesis(ph, ncol)
        else:
            maps = plan.synthesis_lmajor(buf, ncol)
        self.base.n_SHTs_reverse += ncol
        self.timers['map_transforms'] += time.time() - t0
        return maps.view(len(hs), 2, nr, self.base.Npix)

    def _gnl_fused_ok(self):
        """Whether the gNL-loc t4 sum can skip the maps (ps_sht_ring_gnl): gNL-loc alone, an FP32
        plan, and staged ring kernels for both ring classes at this N_r."""
        key = (self.N_r, self.lmax)
        if getattr(self, '_gnl_fused_key', None) != key:
            ok = self.fuse_gnl and self.templates == ['gNL-loc'] and self.base.precision == 32
            if ok:
                ok = self.base.plan(self.lmax, 0).ring_gnl_work_bytes(self.N_r) >= 0
            self._gnl_fused, self._gnl_fused_key = ok, key
        return self._gnl_fused

    def _filter_gnl(self, h, c4, c2=0.0, P_weight=1. / np.sqrt(4. * np.pi)):
        """The gNL-loc t4 sum (tspec_template.py:1391, plus the collapsed disconnected term when
        c2 != 0) of one filtered field as an FP64 device scalar, without its maps: the P and Q
        weightings interleaved as (P_r, Q_r), the Legendre stage, then the ring stage fused with
        the pixel sum."""
        t0 = time.time()
        nr = self.N_r
        ncol = 2 * nr
        buf, plan = self._lmajor_buffer(ncol, self.lmax)
        fPb, fQb, inv_s2 = self._gnl_balanced_tables(P_weight)
        K.weight_alm_pq_lmajor(fPb, fQb, h, self.lmax, self.lmin, P_weight, 1., buf, interleaved=True)
        ph = self._lmaj.get(('phase', ncol))
        if ph is None:
            ph = plan.phase_buffer(ncol)
            self._lmaj[('phase', ncol)] = ph
        if self.kernel_events is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.legendre_synthesis_lmajor(buf, ncol, ph)
            e1.record()
            self.kernel_events.append((ncol, e0, e1))
        else:
            plan.legendre_synthesis_lmajor(buf, ncol, ph)
        ws = self._lmaj.get(('gnl_parts', nr))
        if ws is None:
            ws = torch.empty(plan.ring_gnl_work_bytes(nr), dtype=torch.uint8, device=self.base.device)
            self._lmaj[('gnl_parts', nr)] = ws
        U, V = (self.U_d, self.V_d) if c2 != 0 else (None, None)
        out = plan.ring_gnl(ph, nr, self.w_d['gNL-loc'], c4, U, V, c2, ws=ws, inv_s2=inv_s2)
        self.base.n_SHTs_reverse += ncol
        self.timers['map_transforms'] += time.time() - t0
        return out

    # ---------------------------------------------------------- tauNL pieces
    def _A_maps(self, P1, Q1, P2=None, Q2=None, out=None):
        """[P Q]_LM of fields 1 = (P1, Q1) and 2 = (P2, Q2) (tspec_template.py:1294-1303, n = 0):
        Y^dag[Q1 P1] without field 2, else Y^dag[Q2 P1 + Q1 P2], at nLmax, as (1, N_r, nlm(nLmax))."""
        t0 = time.time()
        if P2 is None:
            prod = K.multiplyRR(Q1, P1)
        else:
            prod = K.multiplyRR_sym(Q2, P1, Q1, P2)
        plan = self.base.plan(self.nLmax, 0)
        A = plan.adjoint(prod.view(self.N_r, 1, -1), np.pi / (3. * self.base.Nside ** 2), out=out)
        self.base.n_SHTs_forward += self.N_r
        self.timers['tauNL_products'] += time.time() - t0
        return A.view(1, self.N_r, -1)

    def _tau_sum(self, A, B, t, out=None, accumulate=False, scale=1.0):
        """tau_sum_n0 (tauNL_utils.pyx:604-646) on device, result scaled on host by `scale`."""
        t0 = time.time()
        r = K.tau_sum_n0(A, B, self.FLL_d, self.w_d[t], self.nLmax, self.Lmin, self.Lmax, 0)
        self.timers['tauNL_summation'] += time.time() - t0
        return r * scale

    # ------------------------------------------------------------------ sims
    def _process_pairs(self, pairs, input_type, with_stack=False):
        """applySinv + filters for a list of (alpha, beta) inputs -> list of (maps_a, maps_b); with
        with_stack also the (2 n, 2, N_r, Npix) stack [a0, b0, a1, b1, ...] and the filtered
        harmonics h of each map."""
        t0 = time.time()
        hs = []
        for a, b in pairs:
            hs.append(self._h_dev(self.applySinv(a, input_type=input_type, lmax=self.lmax)))
            hs.append(self._h_dev(self.applySinv(b, input_type=input_type, lmax=self.lmax)))
        self.timers['Sinv'] += time.time() - t0
        maps = self._filter(hs)
        out = [(maps[2 * i], maps[2 * i + 1]) for i in range(len(pairs))]
        return (out, maps, hs) if with_stack else out

    def _sim_pair(self, ii, Cl_input):
        if self.ones_mask:
            a = self.base.generate_data(int(1e5) + ii, Cl_input=Cl_input, output_type='harmonic', lmax=self.lmax,
                                        deconvolve_beam=False)
            b = self.base.generate_data(int(2e5) + ii, Cl_input=Cl_input, output_type='harmonic', lmax=self.lmax,
                                        deconvolve_beam=False)
            return (a, b), 'harmonic'
        a = self._masked(self.base.generate_data(int(1e5) + ii, Cl_input=Cl_input, deconvolve_beam=False))
        b = self._masked(self.base.generate_data(int(2e5) + ii, Cl_input=Cl_input, deconvolve_beam=False))
        return (a, b), 'map'

    def _beam_times(self, a, lmax=None):
        """beam_lm * a for alm at lmax (default: base lmax), host or device arrays."""
        b = np.asarray(self.base.beam_lm)
        if lmax is not None:
            b = b[:, self.base.l_arr <= lmax]
        if isinstance(a, torch.Tensor):
            return a * torch.from_numpy(np.ascontiguousarray(b)).to(a.device)
        return b * a

    def _masked(self, m):
        if isinstance(m, torch.Tensor):
            if not hasattr(self, 'mask_d'):
                mk = np.asarray(self.mask, dtype=np.float64)
                self.mask_d = torch.from_numpy(mk).to(self.base.device, m.dtype)
            return (m * self.mask_d.to(m.dtype)).contiguous()
        return self.mask * m

    def _collapse_init(self):
        dev = self.base.device
        self.t0_num = np.zeros(len(self.templates))
        # Monte-Carlo sums U = sum_s P_s Q_s, V = sum_s P_s^2 in FP64 whatever the working precision
        # (gNL-loc uses both; tauNL-loc uses U: sum_s A_ss = Y^dag[sum_s Q_s P_s] = Y^dag[U])
        self.U_d = torch.zeros((self.N_r, self.base.Npix), dtype=torch.float64, device=dev)
        self.V_d = torch.zeros((self.N_r, self.base.Npix), dtype=torch.float64, device=dev)
        self._t0_acc = torch.zeros(len(self.templates), dtype=torch.float64, device=dev)
        # sum over the simulations of |P_r|^2 and |Q_r|^2 per node: the magnitude balance of the
        # fused gNL-loc sum's column pairs (_gnl_balanced_tables), measured on maps with the data's
        # statistics
        self._pq_pow = torch.zeros((2, self.N_r), dtype=torch.float64, device=dev)
        self._gnl_tables = None
        # tauNL-loc mean-field shift M (l-major (nlm(nLmax), N_r)): every self-product map A_xx
        # (A_dd, A_aa, A_bb, and sum_s A_ss -> sum_s A_ss - 2 N_it M) enters the tau sums as A_xx - M.
        # For any fixed M the tau(M, .) terms of t4, t2 and t0 cancel exactly (tau is bilinear and
        # the prefactors are 12, -12 / N_it and 12 / N_it per pair); with M near E[A_ss] -- on a
        # masked sky a mean field far larger than the fluctuations -- each term stays near the size
        # of the result, so FP32 A maps lose no more than FP32 accuracy of the result.  M is the
        # mean of A_ss over the first chunk of simulations, fixed before any tau sum uses it.
        self._M_t = None

    def _A_batch_lmajor(self, prods):
        """Y^dag at nLmax of a (n, N_r, Npix) stack of product maps -> l-major (nlm(nLmax), n N_r),
        column s N_r + r (the layout of ps_tau_sum_lmajor)."""
        t0 = time.time()
        n = prods.shape[0]
        plan = self.base.plan(self.nLmax, 0)
        A = plan.adjoint(prods.view(n * self.N_r, 1, -1), np.pi / (3. * self.base.Nside ** 2))
        self.base.n_SHTs_forward += n * self.N_r
        At = plan.to_lmajor(A)
        self.timers['tauNL_products'] += time.time() - t0
        return At

    def _tau_lmajor(self, At, Bt, nmu, t, out, scale=1., shift=(False, False)):
        """out += scale * sum_mu tau_sum_n0(A_mu, B_mu) (device scalar); shift = (on A, on B)
        subtracts the mean-field map M from that side."""
        t0 = time.time()
        kw = dict(M=self._M_t, shiftA=shift[0], shiftB=shift[1]) if any(shift) else {}
        if scale == 1.:
            K.tau_sum_lmajor(At, Bt, self.FLL_d, self.w_d[t], self.N_r, nmu, self.nLmax, self.Lmin, self.Lmax,
                             out=out, accumulate=True, **kw)
        else:
            r = K.tau_sum_lmajor(At, Bt, self.FLL_d, self.w_d[t], self.N_r, nmu, self.nLmax, self.Lmin, self.Lmax,
                                 **kw)
            out.add_(r, alpha=scale)
        self.timers['tauNL_summation'] += time.time() - t0

    def _collapse_add(self, stack):
        """Add a chunk of simulation pairs, stack (2 n, 2, N_r, Npix) = [a0, b0, a1, b1, ...], to
        the Monte-Carlo sums and to t0 (device accumulators; no host synchronisation)."""
        n = stack.shape[0] // 2
        P = stack[:, 0].contiguous()
        Q = stack[:, 1].contiguous()
        K.mc_accumulate(P, Q, self.U_d, self.V_d)
        self._pq_pow[0] += torch.linalg.vector_norm(P, dim=(0, 2), dtype=torch.float64) ** 2
        self._pq_pow[1] += torch.linalg.vector_norm(Q, dim=(0, 2), dtype=torch.float64) ** 2
        for ii, t in enumerate(self.templates):
            if t == 'gNL-loc':
                for i in range(n):
                    K.gnl_loc_disc_sum(self.w_d[t], P[2 * i], P[2 * i + 1], Q[2 * i], Q[2 * i + 1],
                                       out=self._t0_acc[ii:ii + 1].view(()), accumulate=True)
            if t == 'tauNL-loc':
                # t0 (tspec_template.py:1680-1687): sum over pairs of 4 tau(A_aa, A_bb) + 2 tau(A_ab, A_ab),
                # A_xy = Y^dag[Q_x P_y (+ Q_y P_x)], as l-major batches over the pairs
                Pa, Pb = P[0::2].contiguous(), P[1::2].contiguous()
                Qa, Qb = Q[0::2].contiguous(), Q[1::2].contiguous()
                A_aa = self._A_batch_lmajor(K.multiplyRR(Qa, Pa))
                A_bb = self._A_batch_lmajor(K.multiplyRR(Qb, Pb))
                if self._M_t is None:
                    nlm = A_aa.shape[0]
                    c128 = torch.complex128
                    self._M_t = ((A_aa.view(nlm, n, self.N_r).to(c128).sum(1) +
                                  A_bb.view(nlm, n, self.N_r).to(c128).sum(1)) / (2 * n)).to(A_aa.dtype)
                tau = self._t0_acc[ii:ii + 1].view(())
                self._tau_lmajor(A_aa, A_bb, n, t, tau, scale=4., shift=(True, True))
                del A_aa, A_bb
                A_ab = self._A_batch_lmajor(K.multiplyRR_sym(Qb, Pa, Qa, Pb))
                self._tau_lmajor(A_ab, A_ab, n, t, tau, scale=2.)

    def _collapse_finish(self):
        pref = np.zeros(len(self.templates))
        for ii, t in enumerate(self.templates):
            if t == 'gNL-loc':
                pref[ii] = 54. / 100. / self.N_it * self.base.A_pix * FOURPI15
            if t == 'tauNL-loc':
                pref[ii] = 1. / 24. * 3 / self.N_it * FOURPI15
        self.t0_num = self._t0_acc.cpu().numpy() * pref
        del self._t0_acc
        if 'gNL-loc' in self.templates:
            # balance of the fused sum's (P_r, Q_r) column pairs from the simulations' own maps:
            # s_r^2 = rms(Q_r) / rms(P_r)
            pp, qq = self._pq_pow[0], self._pq_pow[1]
            ok = (pp > 0) & (qq > 0)
            s2 = torch.where(ok, torch.sqrt(torch.where(ok, qq, 1.) / torch.where(ok, pp, 1.)), torch.ones_like(pp))
            sr = torch.sqrt(s2)
            rdt = self.base.rdtype
            self._gnl_tables = ((self.fP_d.double() * sr).to(rdt).contiguous(),
                                (self.fQ_d.double() / sr).to(rdt).contiguous(), (1. / s2).contiguous())
        del self._pq_pow
        if 'tauNL-loc' in self.templates:
            # sum_s A_ss = Y^dag[U], l-major (one column per radial node), shifted: D = sum_s (A_ss - M),
            # formed in FP64 (2 N_it M is exact there for an FP32 M)
            SA = self._A_batch_lmajor(self.U_d.to(self.base.rdtype)[None])
            c128 = torch.complex128
            self.D_t = (SA.to(c128) - (2 * self.N_it) * self._M_t.to(c128)).to(SA.dtype)
            del SA
        if 'gNL-loc' not in self.templates:
            del self.U_d, self.V_d

    # ------------------------------------------- simulation store (tauNL-loc)
    def _sim_store_init(self, n_sims):
        """Room for the filtered P, Q stacks of the simulations that the per-numerator tauNL-loc
        A_sd terms revisit: as many as fit in the memory budget stay resident on the device, the
        others keep their filtered harmonics h (npol x nlm(lmax)) and are re-synthesised."""
        per = 2 * self.N_r * self.base.Npix * torch.tensor([], dtype=self.base.rdtype).element_size()
        if self.sim_memory_gb is None:
            free, _ = torch.cuda.mem_get_info(self.base.device)
            budget = 0.7 * free
        else:
            budget = self.sim_memory_gb * 2 ** 30
        n_res = int(min(n_sims, budget // per))
        dev, rdt = self.base.device, self.base.rdtype
        self._sim_P = torch.empty((n_res, self.N_r, self.base.Npix), dtype=rdt, device=dev)
        self._sim_Q = torch.empty((n_res, self.N_r, self.base.Npix), dtype=rdt, device=dev)
        self._sim_h = [None] * n_sims
        self._sim_count = 0

    def _sim_store_add(self, stack, hs):
        n_res = self._sim_P.shape[0]
        for i in range(stack.shape[0]):
            s = self._sim_count
            if s < n_res:
                self._sim_P[s].copy_(stack[i, 0])
                self._sim_Q[s].copy_(stack[i, 1])
            else:
                self._sim_h[s] = hs[i]
            self._sim_count += 1

    def _sim_chunks(self, nb):
        """(P, Q) stacks (n, N_r, Npix) of consecutive simulations, n <= nb, resident or re-made."""
        n_res, n_all = self._sim_P.shape[0], len(self._sim_h)
        for i0 in range(0, n_res, nb):
            i1 = min(n_res, i0 + nb)
            yield self._sim_P[i0:i1], self._sim_Q[i0:i1]
        for i0 in range(n_res, n_all, nb):
            i1 = min(n_all, i0 + nb)
            maps = self._filter(self._sim_h[i0:i1])
            yield maps[:, 0].contiguous(), maps[:, 1].contiguous()

    def _ingest_sims(self, get_pair, N_pairs, preload, verb):
        self._sims_version += 1  # invalidates a captured numerator graph
        self.N_it = N_pairs
        print("Using %d pairs of Monte Carlo simulations" % self.N_it)
        keep_maps = preload and not self.batch_sims
        store_tau = self.batch_sims and 'tauNL-loc' in self.templates
        self.preload = preload
        if self.batch_sims:
            self._collapse_init()
        if store_tau:
            self._sim_store_init(2 * N_pairs)
        if keep_maps:
            self.proc_maps = []
        nb = self.sim_batch or max(1, min(4, int(64 // max(1, self.N_r // 4))))
        process_now = self.batch_sims or keep_maps
        for i0 in range(0, N_pairs if process_now else 0, nb):
            idx = list(range(i0, min(N_pairs, i0 + nb)))
            if verb:
                print("Processing bias simulation pairs %d-%d of %d" % (idx[0] + 1, idx[-1] + 1, N_pairs))
            pairs, itype = [], None
            for ii in idx:
                pr, itype = get_pair(ii)
                pairs.append(pr)
            procs, stack, hs = self._process_pairs(pairs, itype, with_stack=True)
            if self.batch_sims:
                self._collapse_add(stack)
            if store_tau:
                self._sim_store_add(stack, hs)
            if keep_maps:
                for ma, mb in procs:
                    self.proc_maps.append((ma.clone(), mb.clone()))
            del procs, stack, hs
        if self.batch_sims:
            self._collapse_finish()
        if not preload:
            def load_sim_data(ii):
                pr, itype = get_pair(ii)
                return self._process_pairs([pr], itype)[0]
            self.load_sim_data = load_sim_data

    def generate_sims(self, N_pairs, Cl_input=[], preload=True, verb=False):
        """Gaussian simulation pairs for the two- and zero-field terms (tspec_template.py:1172-1228)."""
        if len(Cl_input) == 0:
            Cl_input = self.base.Cl_tot
        self._ingest_sims(lambda ii: self._sim_pair(ii, Cl_input), N_pairs, preload, verb)

    def load_sims(self, load_sim_pair, N_pairs, verb=False, preload=True, input_type='map'):
        """User-supplied simulation pairs (tspec_template.py:1128-1170)."""
        self._ingest_sims(lambda ii: (load_sim_pair(ii), input_type), N_pairs, preload, verb)

    def _sim_maps(self, isim):
        if self.preload and hasattr(self, 'proc_maps'):
            return self.proc_maps[isim]
        return self.load_sim_data(isim)

    # ------------------------------------------------------------ numerator
    def Tl_numerator(self, data, include_disconnected_term=True, verb=False, input_type='map'):
        """Numerator of the unwindowed estimator for every template (tspec_template.py:1338-1954)."""
        t_start = time.time()
        if not hasattr(self, 'r_arr'):
            raise Exception("Need to supply radial integration points or run optimize_radial_sampling_1d()!")
        if not hasattr(self, 'preload') and include_disconnected_term:
            raise Exception("Need to generate or specify bias simulations!")
        if self.pol:
            assert len(data) == 3, "Data must contain T, Q, U components!"
        else:
            assert (len(data) == 1 and len(data[0]) == self.base.Npix) or len(data) == self.base.Npix, \
                "Data must contain T only!"
        compute_t0 = not (hasattr(self, 't0_num') and include_disconnected_term)

        nT = len(self.templates)
        t2 = [0.] * nT
        collapsed = include_disconnected_term and self.batch_sims
        graph_ok = (self.capture_graph and collapsed and input_type == 'map' and isinstance(data, torch.Tensor)
                    and data.is_cuda)
        if graph_ok and not getattr(self.applySinv, 'capturable', False):
            raise Exception("capture_graph needs an applySinv without host synchronisation "
                            "(Weightings.applySinv_ideal); %r is not marked capturable" % (self.applySinv,))
        # the maps P, Q are read again only by the per-simulation disconnected terms and tauNL-loc
        need_maps = (include_disconnected_term and not collapsed) or 'tauNL-loc' in self.templates
        if graph_ok:
            t4_d, P, Q, A_dd = self._numerator_graph(data)
        else:
            t4_d, P, Q, A_dd = self._numerator_device(data, input_type, collapsed, need_maps=need_maps)
        t4 = t4_d.cpu().numpy()

        if include_disconnected_term:
            t2 = np.zeros(nT)
            t0n = np.zeros(nT)
            if collapsed:
                # remaining per-simulation work: tauNL-loc A_sd terms (quadratic in the simulation)
                if 'tauNL-loc' in self.templates:
                    ii = self.templates.index('tauNL-loc')
                    t = 'tauNL-loc'
                    # sum_s 2 tau(A_sd, A_sd), A_sd = Y^dag[Q_s P + Q P_s], in batches of simulations
                    acc = torch.zeros((), dtype=torch.float64, device=self.base.device)
                    nb = self.sim_batch_tau or max(1, 256 // self.N_r)
                    for Ps, Qs in self._sim_chunks(nb):
                        A_sd = self._A_batch_lmajor(K.multiplyRR_sym_batch(Qs, P, Ps, Q))
                        self._tau_lmajor(A_sd, A_sd, Ps.shape[0], t, acc)
                        del A_sd
                    t2[ii] = -1. / 24. * 2. * float(acc) * 6 / self.N_it / 2. * FOURPI15
            else:
                for isim in range(self.N_it):
                    ma, mb = self._sim_maps(isim)
                    for ii, t in enumerate(self.templates):
                        if t == 'gNL-loc':
                            s = (float(K.gnl_loc_disc_sum(self.w_d[t], P, ma[0], Q, ma[1])) +
                                 float(K.gnl_loc_disc_sum(self.w_d[t], P, mb[0], Q, mb[1])))
                            t2[ii] += -54. / 100. * s / self.N_it * self.base.A_pix * FOURPI15
                            if compute_t0:
                                s = float(K.gnl_loc_disc_sum(self.w_d[t], ma[0], mb[0], ma[1], mb[1]))
                                t0n[ii] += 54. / 100. * s / self.N_it * self.base.A_pix * FOURPI15
                        if t == 'tauNL-loc':
                            A_aa = self._A_maps(ma[0], ma[1])
                            A_ad = self._A_maps(ma[0], ma[1], P, Q)
                            tau = 4 * float(self._tau_sum(A_aa, A_dd, t)) + 2 * float(self._tau_sum(A_ad, A_ad, t))
                            A_bb = self._A_maps(mb[0], mb[1])
                            A_bd = self._A_maps(mb[0], mb[1], P, Q)
                            tau += 4 * float(self._tau_sum(A_bb, A_dd, t)) + 2 * float(self._tau_sum(A_bd, A_bd, t))
                            t2[ii] += -1. / 24. * tau * 6 / self.N_it / 2. * FOURPI15
                            if compute_t0:
                                A_ab = self._A_maps(ma[0], ma[1], mb[0], mb[1])
                                tau = 4 * float(self._tau_sum(A_aa, A_bb, t)) + 2 * float(self._tau_sum(A_ab, A_ab, t))
                                t0n[ii] += 1. / 24. * tau * 3 / self.N_it * FOURPI15
            if compute_t0 and not collapsed:
                self.t0_num = t0n
            out = t4 + t2 + self.t0_num
        else:
            print("## No subtraction of (parity-conserving) disconnected terms performed!")
            out = t4
        self.timers['numerator'] += time.time() - t_start
        return out

    def _numerator_device(self, data, input_type, collapsed, need_maps=True):
        """The data-dependent device work of a numerator: S^-1, radial filtering, syntheses and
        the t4 sums (with the collapsed Monte-Carlo terms when `collapsed`).  Returns the (nT,)
        FP64 tensor of t4 values and the maps P, Q and A_dd (tauNL-loc) that the remaining
        terms use (None when `need_maps` is False and the fused gNL-loc sum ran); no host
        synchronisation."""
        t0 = time.time()
        h = self._h_dev(self.applySinv(data, input_type=input_type, lmax=self.lmax))
        self.timers['Sinv'] += time.time() - t0
        if not need_maps and self._gnl_fused_ok():
            t1 = time.time()
            c4 = 9. / 100. * 4. * self.base.A_pix * FOURPI15
            c2 = -54. / 100. * self.base.A_pix * FOURPI15 / self.N_it if collapsed else 0.
            t4 = self._filter_gnl(h, c4, c2)
            self.timers['gNL_summation'] += time.time() - t1
            return t4.reshape(1), None, None, None
        maps = self._filter([h])[0]
        P, Q = maps[0], maps[1]
        t4 = []
        A_dd = None
        for t in self.templates:
            if t == 'gNL-loc':
                t1 = time.time()
                c4 = 9. / 100. * 4. * self.base.A_pix * FOURPI15
                if collapsed:
                    c2 = -54. / 100. * self.base.A_pix * FOURPI15 / self.N_it
                    t4.append(K.gnl_loc_fused_sum(P, Q, self.w_d[t], c4, self.U_d, self.V_d, c2))
                else:
                    t4.append(K.gnl_loc_fused_sum(P, Q, self.w_d[t], c4))
                self.timers['gNL_summation'] += time.time() - t1
            if t == 'tauNL-loc':
                if collapsed:
                    Ad_t = self._A_batch_lmajor(K.multiplyRR(Q, P)[None])
                    v = torch.zeros((), dtype=torch.float64, device=self.base.device)
                    self._tau_lmajor(Ad_t, Ad_t, 1, t, v, scale=1. / 24. * 12. * FOURPI15, shift=(True, True))
                    # 4 tau(sum_s A_ss, A_dd) of the t2 term, linear in the simulations (both shifted by M)
                    self._tau_lmajor(self.D_t, Ad_t, 1, t, v, scale=-1. / 24. * 6. / self.N_it / 2. * FOURPI15 * 4.,
                                     shift=(False, True))
                else:
                    A_dd = self._A_maps(P, Q)
                    v = self._tau_sum(A_dd, A_dd, t, scale=1. / 24. * 12. * FOURPI15)
                t4.append(v)
        return torch.stack([x.reshape(()) for x in t4]), P, Q, A_dd

    def _numerator_graph(self, data):
        """_numerator_device through a captured HIP graph: captured on the first call for this
        data shape and set of simulations (after one eager warm-up run that sizes every cached
        workspace), then replayed with the data copied into the graph's input buffer."""
        key = (tuple(data.shape), data.dtype, self._sims_version, id(self.applySinv))
        if self._graph is None or self._graph_key != key:
            self._graph = None
            # warm-up: plans, workspaces, scratch
            self._numerator_device(data, 'map', True, need_maps='tauNL-loc' in self.templates)
            torch.cuda.synchronize()
            self._g_in = data.detach().clone()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._g_out = self._numerator_device(self._g_in, 'map', True,
                                                     need_maps='tauNL-loc' in self.templates)
            self._graph, self._graph_key = g, key
        self._g_in.copy_(data)
        self._graph.replay()
        t4, P, Q, A_dd = self._g_out
        return t4.clone(), P, Q, A_dd

    # --------------------------------------------------------------- Fisher
    def _transform_maps(self, maps, f, w, out, scale):
        """sum_r w_r f_l^X(r) Y^dag[map_r]_lm (tspec_template.py:2575-2581), added into out."""
        plan = self.base.plan(self.lmax, 0)
        lm = plan.adjoint(maps.view(self.N_r, 1, -1).contiguous(), np.pi / (3. * self.base.Nside ** 2))
        self.base.n_SHTs_forward += self.N_r
        K.radial_sum(lm.view(self.N_r, -1), w, f, self.lmax, self.lmin, 1 + 2 * self.pol, out=out, scale=scale,
                     accumulate=True)

    def _Q_gnl(self, Pm, Qm, Qs, ii):
        """gNL-loc Q-derivatives (tspec_template.py:2713-2740) into Qs[:, ii]."""
        t = 'gNL-loc'
        w = self.w_d[t]
        s4 = np.sqrt(4. * np.pi)
        P0 = [Pm[0] * s4, Pm[1] * s4]
        fp, fq = self.fP_d, self.fQ_d
        self._transform_maps(K.multiply(P0[0], P0[0], Qm[0]), fp, w, Qs[0, ii], 162. / 25.)
        self._transform_maps(K.multiply(P0[0], P0[0], P0[0]), fq, w, Qs[0, ii], 54. / 25.)
        self._transform_maps(K.multiply(P0[1], P0[1], Qm[1]), fp, w, Qs[1, ii], 162. / 25.)
        self._transform_maps(K.multiply(P0[1], P0[1], P0[1]), fq, w, Qs[1, ii], 54. / 25.)
        self._transform_maps(K.multiply_asym(P0[0], P0[1], Qm[0], Qm[1]), fp, w, Qs[2, ii], 54. / 25.)
        self._transform_maps(K.multiply(P0[0], P0[0], P0[1]), fq, w, Qs[2, ii], 54. / 25.)
        self._transform_maps(K.multiply_asym(P0[1], P0[0], Qm[1], Qm[0]), fp, w, Qs[3, ii], 54. / 25.)
        self._transform_maps(K.multiply(P0[0], P0[1], P0[1]), fq, w, Qs[3, ii], 54. / 25.)

    def _Q_tau(self, Pm, Qm, Qs, ii, weight):
        """tauNL-loc Q-derivatives: _F_PQ (2287-2306) + _get_Q4_perms (2521-2554) with n1 = n3 = n = 0."""
        t = 'tauNL-loc'
        w = self.w_d[t]
        Npix = self.base.Npix
        # F_vecs[i] = to_map_vec(convolve_F_n000(Y^dag[PQ])) for (i1, i2) = (0,0), (1,1), (0,1) sym
        F = []
        for i1, i2 in ((0, 0), (1, 1), (0, 1)):
            if i1 == i2:
                A = self._A_maps(Pm[i1], Qm[i1])
            else:
                # Q[i2] P[i1] + Q[i1] P[i2] (multiplyRR_sym, tspec_template.py:1975)
                A = self._A_maps(Pm[i1], Qm[i1], Pm[i2], Qm[i2])
            buf, plan = self._lmajor_buffer(self.N_r, self.nLmax)
            K.convolve_F_lmajor(A[0], self.FLL_d, w, self.nLmax, self.Lmin, self.Lmax, buf)
            F.append(plan.synthesis_lmajor(buf, self.N_r).view(self.N_r, Npix).clone())
            self.base.n_SHTs_reverse += self.N_r
        # (pq_inds, F_inds, weights) of the four Q4 permutations
        perms = [([0], [0], [12.]), ([1], [1], [12.]), ([1, 0], [0, 2], [4., 4.]), ([0, 1], [1, 2], [4., 4.])]
        for q, (pq, fi, wt) in enumerate(perms):
            # step 1: q_l^X integral of [sum w P F]_lm (compute_productP_real_all, integrate_pq)
            if len(pq) == 1:
                prodP = K.multiplyRR(Pm[pq[0]], F[fi[0]])
                prodQ = K.multiplyRR(Qm[pq[0]], F[fi[0]])
            else:
                prodP = K.multiplyRR_sym(Pm[pq[0]], F[fi[0]], Pm[pq[1]], F[fi[1]])
                prodQ = K.multiplyRR_sym(Qm[pq[0]], F[fi[0]], Qm[pq[1]], F[fi[1]])
            self._transform_maps(prodP, self.fQ_d, w, Qs[q, ii], 0.5 * wt[0] * weight)
            # step 2: p_l^X integral of [sum w Q F]_lm with r weights / sqrt(4 pi)
            self._transform_maps(prodQ, self.fP_d, self.w_tau_sqrt[t], Qs[q, ii], 0.5 * wt[0] * weight)

    def _weight_Q(self, Qs, weighting):
        """_weight_Q_maps (tspec_template.py:2556-2572)."""
        nT = len(self.templates)
        for q in range(4):
            for ii in range(nT):
                if weighting == 'Ainv':
                    full = Qs[q, ii] * torch.from_numpy(self.base.beam[:, self.base.l_arr[self.base.l_arr <= self.lmax]]).to(
                        Qs.device, Qs.dtype)
                    t0 = time.time()
                    if self.ones_mask:
                        r = self.applySinv(full, input_type='harmonic', lmax=self.lmax)
                    else:
                        r = self.applySinv(self._masked(self.base.to_map(full, lmax=self.lmax)), lmax=self.lmax)
                    self.timers['Sinv'] += time.time() - t0
                    if not isinstance(r, torch.Tensor):
                        r = torch.from_numpy(np.asarray(r))
                    r = r.to(Qs.device, Qs.dtype)
                    Qs[q, ii] = r * torch.from_numpy(self.lminfilt).to(Qs.device)
                else:
                    Qs[q, ii] = Qs[q, ii] * self.mw_d

    def compute_fisher_contribution(self, seed, verb=False):
        """Fisher contribution from one pair of GRFs (tspec_template.py:2590-3052)."""
        print("Computing Fisher matrix with seed %d" % seed)
        t0 = time.time()
        a_maps = []
        for ii in range(2):
            if self.ones_mask:
                a_maps.append(self.base.generate_data(seed=seed + int((1 + ii) * 1e9), output_type='harmonic',
                                                      lmax=self.lmax, deconvolve_beam=True))
            else:
                a_maps.append(self.base.generate_data(seed=seed + int((1 + ii) * 1e9), output_type='harmonic',
                                                      deconvolve_beam=True))
        self.timers['fish_grfs'] += time.time() - t0
        return self.fisher_contribution_from_grfs(a_maps, verb=verb)

    def fisher_contribution_from_grfs(self, a_maps, verb=False):
        """Fisher contribution of two given harmonic GRFs a_maps (the body of
        compute_fisher_contribution after generate_data, tspec_template.py:2616-3052)."""
        t_start = time.time()
        nT = len(self.templates)
        nslot = 1 + 2 * self.pol
        nlm = (self.lmax + 1) * (self.lmax + 2) // 2

        def compute_Q4(weighting):
            t0 = time.time()
            if weighting == 'Sinv':
                if self.ones_mask:
                    U = [self.applySinv(self._beam_times(a, self.lmax), input_type='harmonic', lmax=self.lmax)
                         for a in a_maps]
                else:
                    U = [self.applySinv(self._masked(self.base.to_map(self._beam_times(a))), lmax=self.lmax)
                         for a in a_maps]
                self.timers['Sinv'] += time.time() - t0
            else:
                if self.ones_mask:
                    U = [self.base.applyAinv(a, input_type='harmonic', lmax=self.lmax) for a in a_maps]
                else:
                    lf = self.base.l_arr <= self.lmax
                    U = [self.base.applyAinv(a, input_type='harmonic') for a in a_maps]
                    U = [u[:, torch.from_numpy(lf).to(u.device)] if isinstance(u, torch.Tensor) else u[:, lf] for u in U]
                self.timers['Ainv'] += time.time() - t0
            hs = [self._h_dev(u) for u in U]
            maps = self._filter(hs)  # (2, 2, N_r, Npix)
            Pm = [maps[0, 0], maps[1, 0]]
            Qm = [maps[0, 1], maps[1, 1]]
            Qs = torch.zeros((4, nT, nslot, nlm), dtype=self.base.cdtype, device=self.base.device)
            t1 = time.time()
            for ii, t in enumerate(self.templates):
                if t == 'gNL-loc':
                    self._Q_gnl(Pm, Qm, Qs, ii)
                if t == 'tauNL-loc':
                    self._Q_tau(Pm, Qm, Qs, ii, FOURPI15)
            self.timers['fish_deriv'] += time.time() - t1
            t1 = time.time()
            self._weight_Q(Qs, weighting)
            self.timers['fish_weighting'] += time.time() - t1
            return Qs.reshape(4, nT, -1).contiguous()

        Q4_Sinv = compute_Q4('Sinv')
        Q4_Ainv = compute_Q4('Ainv')
        t0 = time.time()
        fish = K.outer_product(Q4_Sinv, Q4_Ainv).cpu().numpy()
        self.timers['fish_outer'] += time.time() - t0
        self.timers['fisher'] += time.time() - t_start
        return fish

    def compute_fisher(self, N_it, verb=False):
        """Monte-Carlo Fisher matrix and its inverse (tspec_template.py:3054-3072)."""
        fish = np.zeros((len(self.templates), len(self.templates)))
        for seed in range(N_it):
            print("Computing Fisher contribution %d of %d" % (seed + 1, N_it))
            fish += self.compute_fisher_contribution(seed, verb=verb * (seed == 0)) / N_it
        self.fish = fish
        self.inv_fish = np.linalg.inv(fish)
        return fish

    def Tl_unwindowed(self, data, fish=[], include_disconnected_term=True, verb=False, input_type='map'):
        """Unwindowed amplitudes F^-1 numerator (tspec_template.py:3074-3108)."""
        if len(fish) != 0:
            self.fish = fish
            self.inv_fish = np.linalg.inv(fish)
        if not hasattr(self, 'inv_fish'):
            raise Exception("Need to compute Fisher matrix first!")
        Tl_num = self.Tl_numerator(data, include_disconnected_term=include_disconnected_term, verb=verb,
                                   input_type=input_type)
        Tl_out = np.matmul(self.inv_fish, Tl_num)
        return {t: Tl_out[i] for i, t in enumerate(self.templates)}
