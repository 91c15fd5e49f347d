"""FP64 oracle of the gNL-loc / tauNL-loc estimator (TEST INFRASTRUCTURE ONLY).

Restates polyspec/tspec_template.py for the local templates with the
reference's own structure -- per-simulation loops, (nl)-filtered arrays,
(nLM, N_r) A-map layout -- on top of the oracle SHTs (sht_oracle.c) and loop
nests (kernels_oracle.c).  Inputs are explicit arrays (alm / maps, radial
tables, simulation pairs, Fisher GRFs) so that the GPU path can be fed
identical inputs.

Citations:
  numerator            tspec_template.py:1338-1954 (gNL-loc 1386-1392, 1562-1580;
                       tauNL-loc 1420-1432, 1665-1687)
  filters              _compute_weighted_maps 799-814, _compute_weighted_gaunt_maps 834-855
  A maps               _A_maps 1294-1303
  Fisher contribution  2590-3052 (gNL-loc 2713-2740, tauNL-loc 2842-2848)
  F_PQ / Q4            _compute_PQ_vecs 1967-1980, _convolve_F 2155-2162, _F_PQ 2287-2306,
                       _Q4_contributions 2361-2436, _get_Q4_perms 2521-2554
  weighting of Q       _weight_Q_maps 2556-2572 (unit mask: harmonic S^-1; otherwise
                       S^-1 of the masked map mask * to_map(beam Q), 2567-2569)
  masked Fisher GRFs   compute_fisher_contribution 2631-2644: 'Sinv' S^-1(mask to_map(beam a)),
                       'Ainv' applyAinv(a) at base lmax then the l filter
  applySinv (ideal)    weighting.py:186-214
"""
import numpy as np

from . import (adjoint, apply_fl_weights, convolve_F_n000, getlm, gnl_loc_disc_sum, gnl_loc_sum, integrate_pq,
               outer_product, radial_sum, synthesis, tau_sum_n0, to_lm, to_map)

FOURPI15 = (4. * np.pi) ** 1.5


class OracleTSpec:
    def __init__(self, nside, lmin, lmax, Lmin, Lmax, Cl_tot, beam, pol, templates, plLX0, qlXs, FLL, r_weights,
                 mask=None):
        self.nside, self.lmin, self.lmax, self.Lmin, self.Lmax = nside, lmin, lmax, Lmin, Lmax
        # mask None (or all ones): the unit-mask branches of the reference (ones_mask,
        # tspec_template.py:153); otherwise the masked-sky branches
        m = None if mask is None else np.asarray(mask, dtype=np.float64)
        self.mask = None if m is None or (np.std(m) < 1e-12 and abs(np.mean(m) - 1) < 1e-12) else m
        self.pol = pol
        self.templates = list(templates)
        self.Npix = 12 * nside ** 2
        self.A_pix = 4 * np.pi / self.Npix
        self.base_lmax = 3 * nside - 1
        self.l_arr, self.m_arr = getlm(self.base_lmax)
        self.plLX0 = np.ascontiguousarray(plLX0)
        self.qlXs = np.ascontiguousarray(qlXs)
        self.FLL = None if FLL is None else np.ascontiguousarray(FLL)
        self.r_weights = {k: np.asarray(v, dtype=np.float64) for k, v in r_weights.items()}
        self.N_r = self.qlXs.shape[2]
        # base.py:91-110 inverse total covariance per l
        nf = 3 if pol else 1
        L1 = self.base_lmax + 1
        C = np.zeros((L1, nf, nf))
        names = [['TT', 'TE', 'TB'], ['TE', 'EE', 'EB'], ['TB', 'EB', 'BB']]
        for a in range(nf):
            for b in range(nf):
                k = names[a][b]
                # TB, EB default to zero (base.py:45-46)
                C[:, a, b] = np.asarray(Cl_tot[k])[:L1] if k in Cl_tot else 0.
        inv = np.zeros_like(C)
        good = np.linalg.det(C) > 0
        inv[good] = np.linalg.inv(C[good])
        self.invC = np.transpose(inv, (1, 2, 0))  # (nf, nf, L1)
        self.beam = np.asarray(beam)[:, :L1]
        # l bookkeeping (tspec_template.py:115-125, 437-451)
        self.lfilt = (self.l_arr >= lmin) & (self.l_arr <= lmax)
        self.ls = self.l_arr[self.lfilt].astype(np.int32)
        self.lminfilt = self.l_arr[self.l_arr <= lmax] >= lmin
        self.m_weight = (1. + (self.m_arr > 0))[self.lfilt]
        self.nLmin, self.nLmax = Lmin, Lmax
        nLfilt = (self.l_arr >= Lmin) & (self.l_arr <= Lmax)
        self.Ls = self.l_arr[nLfilt].astype(np.int32)
        self.Ms = self.m_arr[nLfilt].astype(np.int32)
        self.nLminfilt = self.l_arr[self.l_arr <= Lmax] >= Lmin
        # L_indices of tauNL_utils.__init__ (tauNL_utils.pyx:50-62)
        Li = -np.ones(len(self.Ls), dtype=np.int32)
        ip = jp = -1
        for l in self.l_arr:
            if l > Lmax:
                continue
            jp += 1
            if l < Lmin:
                continue
            ip += 1
            if self.Ls[ip] < Lmin or self.Ls[ip] > Lmax:
                continue
            Li[ip] = jp
        self.L_indices = Li

    # -------------------------------------------------------------- S^-1
    def applySinv(self, x, input_type='map', lmax=None):
        """Ideal harmonic S^-1 B C_tot^-1 (weighting.py:196-214), unit mask."""
        lmax = self.lmax if lmax is None else lmax
        lf = self.l_arr <= lmax
        if input_type == 'map':
            alm = to_lm(np.asarray(x), self.nside, self.base_lmax)[:, lf]
        else:
            alm = np.asarray(x)
        ll = self.l_arr[lf]
        Ci = self.invC[:, :, ll]
        return self.beam[:, ll] * np.einsum('abl,bl->al', Ci, alm)

    def applyAinv(self, a, lmax=None):
        """base.py:424-446 (harmonic input)."""
        lmax = self.lmax if lmax is None else lmax
        lf = self.l_arr <= lmax
        ll = self.l_arr[lf]
        b = self.beam[:, ll]
        return b * np.einsum('ijk,jk->ik', self.invC[:, :, ll], b * a)

    # ----------------------------------------------------------- filters
    def _to_map_vec(self, v, lmax):
        return synthesis(np.asarray(v)[:, None, :], self.nside, lmax, 0)[:, 0]

    def _to_lm_vec(self, m, lmax):
        return adjoint(np.asarray(m)[:, None, :], self.nside, lmax, 0)[:, 0] * (np.pi / (3. * self.nside ** 2))

    def weighted_maps(self, h_filt, flX, w):
        summ = np.zeros((flX.shape[2], len(self.lminfilt)), dtype=np.complex128)
        summ[:, self.lminfilt] = apply_fl_weights(flX, h_filt, self.ls, w)
        return self._to_map_vec(summ, self.lmax)

    def filters(self, h_filt):
        """'p00' (1/sqrt(4 pi) weight) and 'q' maps of one filtered field."""
        return {'p00': self.weighted_maps(h_filt, self.plLX0, 1. / np.sqrt(4. * np.pi)),
                'q': self.weighted_maps(h_filt, self.qlXs, 1.)}

    def process(self, x, input_type):
        return self.filters(np.ascontiguousarray(self.applySinv(x, input_type)[:, self.lminfilt]))

    def A_maps(self, m1, m2, add_sym=False):
        if add_sym:
            prod = m2['q'] * m1['p00'] + m1['q'] * m2['p00']
        else:
            prod = m2['q'] * m1['p00']
        out = self._to_lm_vec(prod, self.nLmax).T[self.nLminfilt, :]
        return out[None, :]

    def tau(self, A, B, t):
        return tau_sum_n0(A, B, self.FLL, self.r_weights[t], self.Ls, self.Ms, self.Lmin, self.Lmax, 0)

    # --------------------------------------------------------- numerator
    def numerator(self, data, sims, input_type='map', sim_type='map', include_disconnected_term=True):
        d = self.process(data, input_type)
        N_it = len(sims)
        procs = [(self.process(a, sim_type), self.process(b, sim_type)) for a, b in sims]
        nT = len(self.templates)
        t4, t2, t0 = np.zeros(nT), np.zeros(nT), np.zeros(nT)
        A_dd = None
        for ii, t in enumerate(self.templates):
            if t == 'gNL-loc':
                t4[ii] = 9. / 100. * gnl_loc_sum(self.r_weights[t], d['p00'], d['q']) * 4. * self.A_pix * FOURPI15
            if t == 'tauNL-loc':
                A_dd = self.A_maps(d, d)
                t4[ii] = 1. / 24. * self.tau(A_dd, A_dd, t) * 12. * FOURPI15
        if not include_disconnected_term:
            return t4
        for a, b in procs:
            for ii, t in enumerate(self.templates):
                w = self.r_weights[t]
                if t == 'gNL-loc':
                    perm = lambda m12, m34: gnl_loc_disc_sum(w, m12['p00'], m34['p00'], m12['q'], m34['q'])
                    s = perm(d, a) + perm(d, b)
                    t2[ii] += -54. / 100. * s / N_it * self.A_pix * FOURPI15
                    t0[ii] += 54. / 100. * perm(a, b) / N_it * self.A_pix * FOURPI15
                if t == 'tauNL-loc':
                    A_aa, A_ad = self.A_maps(a, a), self.A_maps(a, d, add_sym=True)
                    A_bb, A_bd = self.A_maps(b, b), self.A_maps(b, d, add_sym=True)
                    tau = 4 * self.tau(A_aa, A_dd, t) + 2 * self.tau(A_ad, A_ad, t)
                    tau += 4 * self.tau(A_bb, A_dd, t) + 2 * self.tau(A_bd, A_bd, t)
                    t2[ii] += -1. / 24. * tau * 6 / N_it / 2. * FOURPI15
                    A_ab = self.A_maps(a, b, add_sym=True)
                    tau = 4 * self.tau(A_aa, A_bb, t) + 2 * self.tau(A_ab, A_ab, t)
                    t0[ii] += 1. / 24. * tau * 3 / N_it * FOURPI15
        return t4 + t2 + t0

    # ------------------------------------------------------------ Fisher
    def _transform(self, m, f, w):
        lm = np.ascontiguousarray(self._to_lm_vec(m, self.lmax)[:, self.lminfilt])
        return radial_sum(lm, w, f, self.ls)

    def fisher_from_grfs(self, a_maps):
        nT = len(self.templates)
        nslot = self.qlXs.shape[1]
        nl = int(np.sum(self.lfilt))

        def compute_Q4(weighting):
            lf = self.l_arr <= self.lmax
            if self.mask is None:  # a_maps at lmax (generate_data(lmax=self.lmax), 2618-2619)
                if weighting == 'Sinv':
                    U = [self.applySinv(self.beam[:, self.l_arr[lf]] * a, 'harmonic')[:, self.lminfilt]
                         for a in a_maps]
                else:
                    U = [self.applyAinv(a)[:, self.lminfilt] for a in a_maps]
            else:  # a_maps at base lmax (2620-2622)
                if weighting == 'Sinv':
                    U = [self.applySinv(self.mask * to_map(self.beam[:, self.l_arr] * a, self.nside, self.base_lmax),
                                        'map')[:, self.lminfilt] for a in a_maps]
                else:
                    U = [self.applyAinv(a, lmax=self.base_lmax)[:, self.lfilt] for a in a_maps]
            Pm = [self.weighted_maps(np.ascontiguousarray(u), self.plLX0, 1. / np.sqrt(4 * np.pi)) for u in U]
            Qm = [self.weighted_maps(np.ascontiguousarray(u), self.qlXs, 1.) for u in U]
            Qs = np.zeros((4, nT, nslot, nl), dtype=np.complex128)
            for ii, t in enumerate(self.templates):
                w = self.r_weights[t]
                if t == 'gNL-loc':
                    P0 = [p * np.sqrt(4. * np.pi) for p in Pm]
                    fp, fq = self.plLX0, self.qlXs
                    Qs[0, ii] = 162. / 25. * self._transform(P0[0] * P0[0] * Qm[0], fp, w)
                    Qs[0, ii] += 54. / 25. * self._transform(P0[0] ** 3, fq, w)
                    Qs[1, ii] = 162. / 25. * self._transform(P0[1] * P0[1] * Qm[1], fp, w)
                    Qs[1, ii] += 54. / 25. * self._transform(P0[1] ** 3, fq, w)
                    Qs[2, ii] = 54. / 25. * self._transform(P0[0] * (2 * P0[1] * Qm[0] + P0[0] * Qm[1]), fp, w)
                    Qs[2, ii] += 54. / 25. * self._transform(P0[0] * P0[0] * P0[1], fq, w)
                    Qs[3, ii] = 54. / 25. * self._transform(P0[1] * (2 * P0[0] * Qm[1] + P0[1] * Qm[0]), fp, w)
                    Qs[3, ii] += 54. / 25. * self._transform(P0[0] * P0[1] * P0[1], fq, w)
                if t == 'tauNL-loc':
                    F = []
                    for i1, i2 in ((0, 0), (1, 1), (0, 1)):
                        if i1 == i2:
                            prod = Qm[i2] * Pm[i1]
                        else:
                            prod = Qm[i2] * Pm[i1] + Qm[i1] * Pm[i2]
                        PQ = np.ascontiguousarray(self._to_lm_vec(prod, self.nLmax).T[self.nLminfilt, :])
                        conv = convolve_F_n000(PQ, self.FLL, w, self.Ls, self.L_indices, len(self.nLminfilt))
                        F.append(self._to_map_vec(conv.T, self.nLmax))
                    perms = [([0], [0], [12.]), ([1], [1], [12.]), ([1, 0], [0, 2], [4., 4.]),
                             ([0, 1], [1, 2], [4., 4.])]
                    for q, (pq, fi, wt) in enumerate(perms):
                        out = np.zeros((nslot, nl), dtype=np.complex128)
                        prodP = sum(wt[k] * Pm[pq[k]] * F[fi[k]] for k in range(len(pq)))
                        PF = np.ascontiguousarray(self._to_lm_vec(prodP, self.lmax).T[self.lminfilt])
                        integrate_pq(self.qlXs, PF, self.ls, w, out)
                        prodQ = sum(wt[k] * Qm[pq[k]] * F[fi[k]] for k in range(len(pq)))
                        QF = np.ascontiguousarray(self._to_lm_vec(prodQ, self.lmax).T[self.lminfilt])
                        integrate_pq(self.plLX0, QF, self.ls, w / np.sqrt(4. * np.pi), out)
                        Qs[q, ii] += FOURPI15 * out
            for q in range(4):
                for ii in range(nT):
                    if weighting == 'Ainv':
                        full = np.zeros((nslot, len(self.lminfilt)), dtype=np.complex128)
                        full[:, self.lminfilt] = self.beam[:, self.ls] * Qs[q, ii]
                        if self.mask is None:
                            Qs[q, ii] = self.applySinv(full, 'harmonic')[:, self.lminfilt]
                        else:
                            Qs[q, ii] = self.applySinv(self.mask * to_map(full, self.nside, self.lmax),
                                                       'map')[:, self.lminfilt]
                    else:
                        Qs[q, ii] = self.m_weight * Qs[q, ii]
            return Qs.reshape(4, nT, -1)

        return outer_product(compute_Q4('Sinv'), compute_Q4('Ainv'))
