"""FP64 oracle of the binned-bispectrum numerator (TEST INFRASTRUCTURE ONLY).

Restates polyspec/bspec_bin.py with the oracle SHTs:
  bins / triangle condition / symmetry factors   bspec_bin.py:92-112, 126-175
  H maps via base.to_map_spin                    bspec_bin.py:177-184, 335-337; base.py:260-297
  cubic and linear numerator terms               bspec_bin.py:317-420
The H maps are built here exactly as the reference builds them: to_map_spin's
(G, C) = (-(a_+ + (-1)^s a_-)/2, -(a_+ - (-1)^s a_-)/(2i)) with the inputs of
compute_spin_transform_map, at the reference's base lmax.
"""
import numpy as np

from . import getlm, synthesis


def to_map_spin(lm_plus, lm_minus, spin, nside, lmax):
    """base.py:260-288 (ducc branch): returns (map_plus, map_minus)."""
    s = (-1) ** spin
    G = -(lm_plus + s * lm_minus) / 2.
    C = -(lm_plus - s * lm_minus) / 2j
    RI = synthesis(np.stack([G, C])[None], nside, lmax, spin)[0]
    return RI[0] + 1j * RI[1], RI[0] - 1j * RI[1]


def spin_transform_map(a_lms, spin, nside, lmax):
    """base.py:290-297: sum_lm sY_lm a_lm for each scalar a_lm."""
    if spin > 0:
        return [to_map_spin(a, (-1.) ** spin * a, abs(spin), nside, lmax)[0] for a in a_lms]
    return [to_map_spin((-1.) ** spin * a, a, abs(spin), nside, lmax)[1] for a in a_lms]


class OracleBSpec:
    def __init__(self, nside, lmax, l_bins, fields, parity, pol):
        self.nside, self.lmax, self.pol = nside, lmax, pol
        self.l_bins = np.asarray(l_bins)
        self.Nl = len(l_bins) - 1
        self.fields = fields
        self.chi_arr = [1] if parity == 'even' else [-1] if parity == 'odd' else [1, -1]
        self.A_pix = 4 * np.pi / (12 * nside ** 2)
        ls, _ = getlm(lmax)
        self.ell_bins = [(ls >= self.l_bins[b]) & (ls < self.l_bins[b + 1]) for b in range(self.Nl)]
        ind = {'T': 0, 'E': 1, 'B': 2}
        par = {'T': 1, 'E': 1, 'B': -1}
        self.all_bins = []
        for u in fields:
            u1, u2, u3 = [ind[c] for c in u]
            p_u = par[u[0]] * par[u[1]] * par[u[2]]
            for b1 in range(self.Nl):
                for b2 in range(self.Nl):
                    if u1 == u2 and b2 < b1:
                        continue
                    for b3 in range(self.Nl):
                        if u2 == u3 and b3 < b2:
                            continue
                        c = [0.5 * (self.l_bins[b] + self.l_bins[b + 1]) for b in (b1, b2, b3)]
                        if c[2] < abs(c[0] - c[1]) or c[2] > c[0] + c[1]:
                            continue
                        self.all_bins.append((u1, u2, u3, b1, b2, b3, p_u))
        sym = []
        for u1, u2, u3, b1, b2, b3, _ in self.all_bins:
            if b1 == b2 and u1 == u2 and b2 == b3 and u2 == u3:
                sym.append(6)
            elif (b1 == b2 and u1 == u2) or (b2 == b3 and u2 == u3):
                sym.append(2)
            else:
                sym.append(1)
        self.sym = np.asarray(sym * len(self.chi_arr))

    def H_maps(self, h):
        """bspec_bin.py:335-337: p1[bin][field], m2[bin][field] complex maps."""
        p1 = [spin_transform_map(self.ell_bins[b] * h.conj(), +1, self.nside, self.lmax) for b in range(self.Nl)]
        m2 = [spin_transform_map(self.ell_bins[b] * h.conj(), -2, self.nside, self.lmax) for b in range(self.Nl)]
        return p1, m2

    def _sum(self, terms):
        out = []
        for chi in self.chi_arr:
            for i, (u1, u2, u3, b1, b2, b3, p_u) in enumerate(self.all_bins):
                t = terms(u1, u2, u3, b1, b2, b3)
                out.append(np.real(np.sum(t * 1j if p_u * chi == -1 else t)))
        return np.asarray(out)

    def numerator(self, h_data, h_sims=()):
        """Bl_numerator (bspec_bin.py:317-420) for S^-1-filtered harmonic data and simulations."""
        p1, m2 = self.H_maps(h_data)
        b3 = self._sum(lambda u1, u2, u3, b1, b2, b3: (p1[b1][u1] * p1[b2][u2] * m2[b3][u3]
                                                        + p1[b2][u2] * p1[b3][u3] * m2[b1][u1]
                                                        + p1[b3][u3] * p1[b1][u1] * m2[b2][u2]))
        b3 = self.A_pix * b3 / 3. / self.sym
        b1 = np.zeros_like(b3)
        for hs in h_sims:
            q1, n2 = self.H_maps(hs)

            def lin(u1, u2, u3, b1_, b2, b3_):
                t = 0
                for (a, b, c) in ((p1, q1, n2), (q1, p1, n2), (q1, q1, m2)):
                    t = t + a[b1_][u1] * b[b2][u2] * c[b3_][u3] + a[b2][u2] * b[b3_][u3] * c[b1_][u1] \
                        + a[b3_][u3] * b[b1_][u1] * c[b2][u2]
                return t
            b1 -= self.A_pix * self._sum(lin) / 3. / self.sym / len(h_sims)
        return b3 + b1
