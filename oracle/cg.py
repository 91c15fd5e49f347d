"""FP64 oracle of the optimal S^-1 filter (TEST INFRASTRUCTURE ONLY).

Restates polyspec/weighting.py with the oracle SHTs:
  Weightings.__init__ noise / spectra set-up   weighting.py:6-85
  applySinv_optimal (x0, CGD, truncation)      weighting.py:148-184
  CGD (_get_b, _apply_A, _apply_M, step, run)  weighting.py:217-367
Vectors are (nf, nlm(base_lmax)) complex128, nf = 1 (T) or 3 (T, E, B).
"""
import numpy as np

from . import getlm, to_lm, to_map


class OracleCG:
    def __init__(self, nside, Cl_raw, Cl_tot, beam, mask, noise_cov, pol):
        self.nside, self.pol = nside, pol
        self.L = 3 * nside - 1
        self.l_arr, self.m_arr = getlm(self.L)
        nf = 3 if pol else 1
        self.nf = nf
        self.A_pix = 4 * np.pi / (12 * nside ** 2)
        self.mask = np.asarray(mask, dtype=np.float64).reshape(nf, -1)
        self.noise_cov = np.asarray(noise_cov, dtype=np.float64).reshape(nf, -1)
        self.noise_icov = np.where(self.noise_cov != 0, 1. / np.where(self.noise_cov != 0, self.noise_cov, 1), 0)
        self.beam = np.asarray(beam)[:, :self.L + 1]
        L1 = self.L + 1
        names = [['TT', 'TE', 'TB'], ['TE', 'EE', 'EB'], ['TB', 'EB', 'BB']]

        def mat(C):
            M = np.zeros((L1, nf, nf))
            for a in range(nf):
                for b in range(nf):
                    k = names[a][b]
                    M[:, a, b] = np.asarray(C[k])[:L1] if k in C else 0.
            return M
        Craw, Ctot = mat(Cl_raw), mat(Cl_tot)
        self.Craw = Craw
        inv = np.zeros_like(Craw)
        if pol and np.sum(np.asarray(Cl_raw['BB'])) == 0:
            good = np.linalg.det(Craw[:, :2, :2]) > 0
            inv[good, :2, :2] = np.linalg.inv(Craw[good, :2, :2])
        else:
            good = np.linalg.det(Craw) > 0
            inv[good] = np.linalg.inv(Craw[good])
        self.invC = inv
        invT = np.zeros_like(Ctot)
        good = np.linalg.det(Ctot) > 0
        invT[good] = np.linalg.inv(Ctot[good])
        self.invCtot = invT
        self.Nl_iso = np.array([np.mean(self.noise_cov[i][self.mask[i] > 0.99]) * self.A_pix for i in range(nf)])
        inv_S = inv + np.einsum('il,jl,ij->lij', self.beam, self.beam, np.diag(1. / self.Nl_iso))
        inv_S[:2] = inv_S[2]
        self.Sigma = np.linalg.inv(inv_S)

    def _lmat(self, M, x):
        return np.einsum('lab,bl->al', M[self.l_arr], x)

    def _beam(self, x):
        return x * self.beam[:, self.l_arr]

    def to_map(self, a):
        return to_map(a, self.nside, self.L)

    def to_lm(self, m):
        return to_lm(m, self.nside, self.L)

    def get_b(self, d):
        return self._beam(self.to_lm(self.mask * d * self.noise_icov / self.A_pix))

    def apply_A(self, x):
        m = self.mask * self.to_map(self._beam(x))
        m = m * self.noise_icov / self.A_pix
        return self._lmat(self.invC, x) + self._beam(self.to_lm(self.mask * m))

    def apply_M(self, r, kind):
        alm = self._lmat(self.Sigma, r)
        if kind == 'harmonic':
            return alm
        sig = self._lmat(self.invC, alm)
        bn = self.beam[:, self.l_arr] / self.Nl_iso[:, None]
        noise = self.to_lm(self.noise_cov * self.to_map(alm * bn)) * self.A_pix * bn
        return self._lmat(self.Sigma, noise + sig)

    def dot(self, a, b):
        return float(np.sum(a * np.conj(b) * (1. + (self.m_arr > 0))).real)

    def solve(self, d, preconditioner='harmonic', nstep=25, thresh=1e-6, x0=True):
        """applySinv_optimal's CG (weighting.py:174-184, 218-351). Returns (ivar_lm, err history)."""
        xs = (np.einsum('lab,bl->al', self.Craw[self.l_arr],
                        self.beam[:, self.l_arr] * self._lmat(self.invCtot, self.to_lm(d)))
              if x0 else np.zeros((self.nf, len(self.l_arr)), complex))
        b0 = self.get_b(d)
        b = b0 - self.apply_A(xs)
        x = np.zeros_like(b)
        r = b.copy()
        z = self.apply_M(r, preconditioner)
        rz = self.dot(r, z)
        rz0 = self.dot(b0, self.apply_M(b0, preconditioner))
        p = z
        errs = []
        for i in range(nstep):
            Ap = self.apply_A(p)
            alpha = rz / self.dot(p, Ap)
            x = x + alpha * p
            r = r - alpha * Ap
            z = self.apply_M(r, preconditioner)
            nrz = self.dot(r, z)
            errs.append(nrz / rz0)
            beta = nrz / rz
            rz = nrz
            p = z + beta * p
            if errs[-1] < thresh:
                break
        return self._lmat(self.invC, x + xs), errs, x + xs
