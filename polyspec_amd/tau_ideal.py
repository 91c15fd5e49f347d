"""tauNL-loc ideal Fisher score of the radial-node optimizer on the MI355X backend.

Mirrors, for n1 = n3 = n = 0 (the local template):
  _compute_fisher_derivatives, tauNL branch     tspec_template.py:3906-3951, 3996-4009, 4062-4068
  _compute_ideal_Q4                              4071-4099
  _get_Q4_perms / _Q4_contributions (n1 = 0)     2521-2554, 2341-2436
  _F_PQ / _compute_PQ_vecs / _convolve_F         2287-2306, 1967-1980, 2142-2162
  _tau_direc_score (score, full score, gradient) 3231-3346, 3485-3486
  apply_ideal_weight, outer_product(_ideal)      ideal_fisher.pyx:499-522, 423-497

The Fisher of the template is F(w) = <Q4(w), W Q4(w)> with W = m_weight C^-1 (no beam: the GRFs
are beam-deconvolved) and the four-permutation combination of outer_product.  Q4(w) is quadratic
in the node weights (the F_L convolution over r' and the radial integral over r both carry w).
The optimizer scores a node subset S with weights w_S by the Fisher distance
<Q4 - Q4_S, W (Q4 - Q4_S)> to the full-grid Q4, averaged over realizations.

Per realization the state holds the P and Q maps of both GRFs and the three [P Q]_LM vectors at
every node of the starting grid; a subset is a gather of rows (the reference instead appends the
maps of each node it adds), so one trial costs 3 F_L convolutions, 3 |S| syntheses at Lmax and
8 |S| adjoints at lmax on the nodes kept.

The weight derivative dQ4[r] follows _Q4_contributions as written: the q-side term uses
del_weights, the p-side term (tspec_template.py:2436) uses r_weights / sqrt(4 pi).  With all
weights equal to the quadrature weights (the Fisher derivative matrix) the two coincide and
sum_r dQ4[r] = Q4; on the gradient path of the weight optimisation (r_weights = del_weights * w)
they differ, as in the reference.
"""
import numpy as np
import torch

from . import kernels as K

FOURPI15 = (4. * np.pi) ** 1.5
# outer_product (ideal_fisher.pyx:423-459): fish = Re sum_pp' C[p, p'] Qa_p^* Qb_p' / (24 * 48)
OUTER_C = np.array([[1., 0., 0., -3.], [0., 1., -3., 0.], [0., -3., 9., 0.], [-3., 0., 0., 9.]])
# _get_Q4_perms (tspec_template.py:2540-2547): (P/Q field indices, F_PQ indices, weight)
PERMS = [([0], [0], 12.), ([1], [1], 12.), ([1, 0], [0, 2], 4.), ([0, 1], [1, 2], 4.)]


def outer_gemm(A, B, chunk_bytes=1 << 30):
    """outer_product (sym=False) of A (4, na, n) and B (4, nb, n) complex as FP64 GEMMs over
    chunks of n: Re sum_pp' C[p, p'] A_p^H B_p' / (24 * 48) -> (na, nb) float64."""
    assert A.dim() == 3 and B.dim() == 3 and A.shape[0] == 4 and B.shape[0] == 4 and A.shape[2] == B.shape[2]
    na, nb, n = A.shape[1], B.shape[1], A.shape[2]
    out = torch.zeros((na, nb), dtype=torch.float64, device=A.device)
    step = max(1, int(chunk_bytes // (16 * (na + nb) * 4)))
    for j0 in range(0, n, step):
        j1 = min(n, j0 + step)
        a = A[:, :, j0:j1].to(torch.complex128)
        b = B[:, :, j0:j1].to(torch.complex128)
        for p in range(4):
            g = torch.zeros_like(b[0])
            for q in range(4):
                if OUTER_C[p, q] != 0.:
                    g.add_(b[q], alpha=OUTER_C[p, q])
            ar = torch.view_as_real(a[p].contiguous()).reshape(na, -1)
            gr = torch.view_as_real(g).reshape(nb, -1)
            out.add_(ar @ gr.T)
    return out / (24. * 48.)


class TauIdeal:
    """The tauNL-loc ideal score of a TSpecTemplate whose tables hold the starting grid."""

    def __init__(self, tm):
        self.tm = tm
        b = tm.base
        self.dev = b.device
        self.nslot = 1 + 2 * tm.pol
        self.nlm = (tm.lmax + 1) * (tm.lmax + 2) // 2
        self.N0 = tm.N_r
        self.Cinv = np.ascontiguousarray(b.inv_Cl_tot_mat[:self.nslot, :self.nslot, :tm.lmax + 1])
        self.quad = np.asarray(tm.quad_weights_1d, dtype=np.float64)
        self.states = []
        self.Q4 = []

    # --------------------------------------------------------------- state
    def add_realization(self, a_maps):
        """P, Q and [P Q]_LM at every node for one pair of beam-deconvolved harmonic GRFs (base
        lmax): U = A^-1 a restricted to l <= lmax (tspec_template.py:3929-3937)."""
        tm = self.tm
        lf = torch.from_numpy(tm.base.l_arr <= tm.lmax).to(self.dev)
        U = []
        for a in a_maps:
            u = tm.base.applyAinv(a, input_type='harmonic')
            u = u if isinstance(u, torch.Tensor) else torch.from_numpy(np.asarray(u))
            U.append(u.to(self.dev)[:, lf])
        maps = tm._filter([tm._h_dev(u) for u in U])
        P = [maps[0, 0].clone(), maps[1, 0].clone()]
        Q = [maps[0, 1].clone(), maps[1, 1].clone()]
        del maps
        # [P Q]_LM of the pairings (0, 0), (1, 1) and (0, 1) + (1, 0) (_F_PQ, 2296-2300)
        PQ = [tm._A_maps(P[0], Q[0])[0].clone(), tm._A_maps(P[1], Q[1])[0].clone(),
              tm._A_maps(P[0], Q[0], P[1], Q[1])[0].clone()]
        st = {'P': P, 'Q': Q, 'PQ': PQ}
        self.states.append(st)
        return st

    # ------------------------------------------------------------ weighting
    def weight(self, x):
        """apply_ideal_weight (ideal_fisher.pyx:499-522): m_weight C^-1 x for x (..., nslot, nlm)."""
        tm = self.tm
        shp = x.shape
        y = K.lmat_apply_batch(x.reshape(-1, self.nslot, self.nlm).contiguous(), tm.lmax, mat=self.Cinv)
        return (y * tm.mw_d.to(y.dtype)).view(shp)

    # ------------------------------------------------------------------ Q4
    def q4(self, st, inds, w, del_w=None):
        """Q4 (4, nslot, nlm) on the nodes inds with weights w, FOURPI15 included
        (_tau_direc_score weight, 3486); with del_w also dQ4 (4, n, nslot, nlm)."""
        tm = self.tm
        inds = np.ascontiguousarray(inds, dtype=np.int64)
        n = len(inds)
        full = n == self.N0 and np.array_equal(inds, np.arange(self.N0))
        idx = torch.from_numpy(inds).to(self.dev)
        take = (lambda x, d: x) if full else (lambda x, d: x.index_select(d, idx).contiguous())
        FLL = take(take(tm.FLL_d, 1), 2)
        fP, fQ = take(tm.fP_d, 2), take(tm.fQ_d, 2)
        P = [take(p, 0) for p in st['P']]
        Q = [take(q, 0) for q in st['Q']]
        w = np.asarray(w, dtype=np.float64)
        wd = torch.from_numpy(w).to(self.dev)
        wd4 = torch.from_numpy(w / np.sqrt(4. * np.pi)).to(self.dev)
        Npix = tm.base.Npix
        # F_L * [P Q] maps at the nodes kept (_convolve_F with inds, 2155-2162)
        F = []
        for k in range(3):
            buf, plan = tm._lmajor_buffer(n, tm.nLmax)
            K.convolve_F_lmajor(take(st['PQ'][k], 0), FLL, wd, tm.nLmax, tm.Lmin, tm.Lmax, buf)
            F.append(plan.synthesis_lmajor(buf, n).view(n, Npix).clone())
            tm.base.n_SHTs_reverse += n
        plan = tm.base.plan(tm.lmax, 0)
        dOmega = np.pi / (3. * tm.base.Nside ** 2)
        Q4 = torch.zeros((4, self.nslot, self.nlm), dtype=tm.base.cdtype, device=self.dev)
        dQ = None
        if del_w is not None:
            dQ = torch.zeros((4, n, self.nslot, self.nlm), dtype=tm.base.cdtype, device=self.dev)
            dd = torch.from_numpy(np.asarray(del_w, dtype=np.float64)).to(self.dev)
        for q, (pq, fi, wt) in enumerate(PERMS):
            if len(pq) == 1:
                prodP = K.multiplyRR(P[pq[0]], F[fi[0]])
                prodQ = K.multiplyRR(Q[pq[0]], F[fi[0]])
            else:
                prodP = K.multiplyRR_sym(P[pq[0]], F[fi[0]], P[pq[1]], F[fi[1]])
                prodQ = K.multiplyRR_sym(Q[pq[0]], F[fi[0]], Q[pq[1]], F[fi[1]])
            sc = 0.5 * wt * FOURPI15
            lmP = plan.adjoint(prodP.view(n, 1, -1), dOmega).view(n, -1)
            del prodP
            K.radial_sum(lmP, wd, fQ, tm.lmax, tm.lmin, self.nslot, out=Q4[q], scale=sc, accumulate=True)
            if dQ is not None:
                K.radial_scatter(lmP, dd, fQ, tm.lmax, tm.lmin, self.nslot, out=dQ[q], scale=sc, accumulate=True)
            del lmP
            lmQ = plan.adjoint(prodQ.view(n, 1, -1), dOmega).view(n, -1)
            del prodQ
            K.radial_sum(lmQ, wd4, fP, tm.lmax, tm.lmin, self.nslot, out=Q4[q], scale=sc, accumulate=True)
            if dQ is not None:
                K.radial_scatter(lmQ, wd4, fP, tm.lmax, tm.lmin, self.nslot, out=dQ[q], scale=sc, accumulate=True)
            del lmQ
            tm.base.n_SHTs_forward += 2 * n
        return Q4, dQ

    # ---------------------------------------------------- Fisher derivative
    def derivative(self, st):
        """One realization of _compute_direc_deriv_single (3950-3973): the full-grid Q4, W Q4, the
        score <Q4, W Q4> and the derivative matrix D[r, r'] = <dQ4[r], W dQ4[r']>."""
        allin = np.arange(self.N0)
        Q4, dQ = self.q4(st, allin, self.quad, del_w=self.quad)
        WQ4 = self.weight(Q4)
        self.Q4.append((Q4, WQ4))
        score = float(K.outer_product(Q4.view(4, 1, -1), WQ4.view(4, 1, -1))[0, 0])
        WdQ = self.weight(dQ)
        D = outer_gemm(dQ.view(4, self.N0, -1), WdQ.view(4, self.N0, -1))
        del dQ, WdQ
        # outer_product_ideal(sym=True) fills the upper triangle and mirrors it
        D = torch.triu(D) + torch.triu(D, 1).T
        return score, D.cpu().numpy()

    # ---------------------------------------------------------------- score
    def score(self, inds, w_vals, scaling=1., gradient=False, full_score=False):
        """_tau_direc_score (3231-3346) for the subset inds with weights quad[inds] w_vals:
        the Fisher distance to the full-grid Q4 averaged over realizations; with gradient also
        its derivative along w_vals, with full_score also the subset's own Fisher."""
        inds = np.asarray(inds, dtype=np.int64)
        this_w = self.quad[inds] * np.asarray(w_vals, dtype=np.float64)
        nf = len(self.states)
        tscore, fscore = 0., 0.
        dscore = np.zeros(len(inds))
        for st, (Q4f, WQ4f) in zip(self.states, self.Q4):
            Q4w, dQ = self.q4(st, inds, this_w, del_w=self.quad[inds] if gradient else None)
            WQ4w = self.weight(Q4w)
            tscore += float(K.outer_product((Q4f - Q4w).view(4, 1, -1), (WQ4f - WQ4w).view(4, 1, -1))[0, 0])
            if full_score:
                fscore += float(K.outer_product(Q4w.view(4, 1, -1), WQ4w.view(4, 1, -1))[0, 0])
            if gradient:
                dscore += 4. * outer_gemm(dQ.view(4, len(inds), -1), (WQ4w - WQ4f).view(4, 1, -1)).cpu().numpy()[:, 0]
        if gradient:
            return tscore / scaling / nf, dscore / scaling / nf
        if full_score:
            return tscore / scaling / nf, fscore / scaling / nf
        return tscore / scaling / nf
