"""S^-1 filters on the MI355X backend: mirrors polyspec/weighting.py.

``Weightings`` (weighting.py:5-214) provides applySinv_ideal and
applySinv_optimal; ``CGD`` (weighting.py:217-367) is the conjugate-gradient
solver of A x = b with A = C^-1 + P^dag N^-1 P, P = W Y B, run entirely on the
device: SHTs through the HIP backend, per-l 3x3 products, m-weighted FP64 dot
products and fused vector updates through libpolyspec_hip.so.  The solver
reproduces the reference's Fletcher-Reeves beta, its x0 start, its operating
lmax (base_lmax = 3 Nside - 1) and its stopping statistic rz/rz0
(weighting.py:261, 336); it also records ||r||/||b|| per iteration.
"""
import numpy as np
import torch

from . import kernels as K


def _per_l_inverse(C, good):
    out = np.zeros_like(C)
    out[good] = np.linalg.inv(C[good])
    return out


class Weightings:
    """S^-1 weighting (weighting.py:5-85).

    Inputs as the reference: base (PolySpec, hip backend), mask (nf, Npix),
    Cl_raw (dict of signal spectra), noise_cov (nf, Npix) pixel variance.
    Inpainting (weighting.py:87-146) is not provided by this backend.
    """

    def __init__(self, base, mask, Cl_raw, noise_cov=None, inpainting_mask=None):
        if inpainting_mask is not None:
            raise NotImplementedError("inpainting is not provided by the hip backend")
        self.base = base
        self.mask = mask
        self.Cl_raw = Cl_raw
        self.pol = self.base.pol
        self.pixel_area = 4. * np.pi / (12 * self.base.Nside ** 2)
        nf = 1 + 2 * self.pol
        self.nf = nf
        L = self.base.lmax
        ls = np.arange(L + 1)
        dev = self.base.device

        m = np.asarray(mask, dtype=np.float64)
        if m.ndim == 1:
            m = np.repeat(m[None], nf, axis=0)
        assert m.shape[-1] == self.base.Npix, f'Mask has incorrect shape: {m.shape}'
        self.mask_arr = m
        self.mask_d = torch.from_numpy(np.ascontiguousarray(m)).to(dev, self.base.rdtype)
        self.beam_lm = self.base.beam_lm
        self.beam_l = np.ascontiguousarray(self.base.beam[:, :L + 1])

        if noise_cov is not None:
            self.noise_cov = np.asarray(noise_cov, dtype=np.float64).reshape(nf, -1)
            self.noise_icov = np.zeros_like(self.noise_cov)
            nz = self.noise_cov != 0
            self.noise_icov[nz] = 1. / self.noise_cov[nz]
            # Isotropic noise level (weighting.py:46-49)
            self.Nl_iso = np.zeros(nf)
            for i in range(nf):
                self.Nl_iso[i] = np.mean(self.noise_cov[i][self.mask_arr[i] > 0.99]) * self.pixel_area
            # C^-1 of the signal spectra, TE block only when B vanishes (weighting.py:51-73)
            C = np.zeros((L + 1, nf, nf))
            if self.pol:
                assert len(Cl_raw.keys()) == 6 or len(Cl_raw.keys()) == 4, \
                    "Must specify four or six input spectra: {ClTT, ClTE, ClEE, ClBB} or {ClTT, ClTE, ClTB, ClEE, ClEB, ClBB}"
                if len(Cl_raw.keys()) == 4:
                    Cl_raw['TB'] = Cl_raw['EB'] = 0. * Cl_raw['TT']
                names = [['TT', 'TE', 'TB'], ['TE', 'EE', 'EB'], ['TB', 'EB', 'BB']]
                for a in range(3):
                    for b in range(3):
                        C[:, a, b] = np.asarray(Cl_raw[names[a][b]])[:L + 1]
                inv = np.zeros_like(C)
                if np.sum(Cl_raw['BB']) == 0:
                    good = np.linalg.det(C[:, :2, :2]) > 0
                    inv[good, :2, :2] = np.linalg.inv(C[good, :2, :2])
                else:
                    good = np.linalg.det(C) > 0
                    inv[good] = np.linalg.inv(C[good])
            else:
                C[:, 0, 0] = np.asarray(Cl_raw['TT'])[:L + 1]
                good = C[:, 0, 0] > 0
                inv = _per_l_inverse(C, good)
            self.Cl_l = C
            self.inv_Cl_l = np.ascontiguousarray(np.transpose(inv, (1, 2, 0)))  # (nf, nf, L+1)
            # Sigma = (C^-1 + B^2/N_iso)^-1, l < 2 copied from l = 2 (weighting.py:75-78)
            inv_S = np.transpose(inv, (1, 2, 0)) + np.einsum('il,jl,ij->ijl', self.beam_l, self.beam_l,
                                                              np.diag(1. / self.Nl_iso))
            inv_S[:, :, ls < 2] = inv_S[:, :, 2:3]
            self.Sigma_l = np.ascontiguousarray(np.transpose(np.linalg.inv(np.transpose(inv_S, (2, 0, 1))), (1, 2, 0)))
            # Device tables
            self.noise_cov_d = torch.from_numpy(self.noise_cov).to(dev, self.base.rdtype)
            wpix = self.mask_arr ** 2 * self.noise_icov / self.pixel_area
            self.wpix_d = torch.from_numpy(np.ascontiguousarray(wpix)).to(dev, self.base.rdtype)
            self.Ninv_d = torch.from_numpy(np.ascontiguousarray(self.mask_arr * self.noise_icov / self.pixel_area)).to(
                dev, self.base.rdtype)
            f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
            self.inv_Cl_l_d = f64(self.inv_Cl_l)
            self.Sigma_l_d = f64(self.Sigma_l)
            self.Cl_l_d = f64(np.transpose(self.Cl_l, (1, 2, 0)))
            self.bn_d = f64(self.beam_l / self.Nl_iso[:, None])
            self.bnA_d = f64(self.beam_l / self.Nl_iso[:, None] * self.pixel_area)
        self.beam_l_d = torch.from_numpy(np.ascontiguousarray(self.beam_l)).to(dev)
        self.invCtot_d = torch.from_numpy(np.ascontiguousarray(self.base.inv_Cl_tot_mat)).to(dev)

    # per-(l,m) views with the reference's attribute names (weighting.py:56-78)
    @property
    def inv_Cl_lm_mat(self):
        return self.inv_Cl_l[:, :, self.base.l_arr]

    @property
    def Sigmalm(self):
        return self.Sigma_l[:, :, self.base.l_arr]

    # ----------------------------------------------------------- filters
    def _dev_vec(self, x):
        if isinstance(x, torch.Tensor):
            return x.to(self.base.device, self.base.cdtype).contiguous()
        return torch.from_numpy(np.ascontiguousarray(x)).to(self.base.device, self.base.cdtype)

    def applySinv_ideal(self, input_map, input_type='map', lmax=None):
        """Harmonic S^-1: B C_tot^-1 Y^dag d (weighting.py:186-214); output (nf, nlm(lmax))."""
        if lmax is None:
            lmax = self.base.lmax
        npy = not isinstance(input_map, torch.Tensor)
        Ltot = self.invCtot_d[:, :, :lmax + 1].contiguous()
        bl = self.beam_l_d[:, :lmax + 1].contiguous()
        if input_type == 'harmonic':
            assert np.mean(self.mask) == 1, "Harmonic-space S^-1 requires a unit mask!"
            a = self._dev_vec(input_map)
            out = K.lmat_apply(a, lmax, mat=Ltot, post=bl)
        elif input_type == 'map':
            m = input_map if not npy else torch.from_numpy(np.ascontiguousarray(input_map))
            # The reference transforms at base_lmax and keeps l <= lmax (weighting.py:210-214);
            # the adjoint coefficients with l <= lmax do not depend on the higher l, so the
            # transform is evaluated at lmax directly.
            alm = self.base.to_lm(m.to(self.base.device, self.base.rdtype).reshape(self.nf, -1).contiguous(),
                                  lmax=lmax)
            out = K.lmat_apply(alm, lmax, mat=Ltot, post=bl)
        else:
            raise AssertionError("input_type must be 'map' or 'harmonic'")
        return self.base._ret(out, npy)

    def applySinv_ideal_batch(self, maps, lmax=None, tune=True):
        """applySinv_ideal(input_type='map') of a stack of map sets, (B, nf, Npix) -> (B, nf, nlm(lmax))
        on the device: one batched adjoint per spin for all B sets (matrix-core kernels from a few
        sets on), then the per-l filter B C_tot^-1 of every set (weighting.py:186-214 on each set).
        tune=False keeps the built-in launch shapes for a batch width not tuned yet (no tuning pass)."""
        if lmax is None:
            lmax = self.base.lmax
        Ltot = self.invCtot_d[:, :, :lmax + 1].contiguous()
        bl = self.beam_l_d[:, :lmax + 1].contiguous()
        m = maps if isinstance(maps, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(maps))
        m = m.to(self.base.device, self.base.rdtype).reshape(m.shape[0], self.nf, -1).contiguous()
        alm = self.base.to_lm_batch(m, lmax=lmax, tune=tune)
        return K.lmat_apply_batch(alm, lmax, mat=Ltot, post=bl)

    def apply_A(self, x):
        """The CG system matrix A x = C^-1 x + B Y^dag W N^-1 W Y B x at base_lmax
        (weighting.py:297-299); x (nf, nlm(base_lmax)) on the device."""
        L = self.base.lmax
        bx = K.lmat_apply(x, L, post=self.beam_l_d)
        m = self.base.to_map(bx)
        m = K.multiplyRR(m, self.wpix_d)
        a = self.base.to_lm(m)
        out = K.lmat_apply(x, L, mat=self.inv_Cl_l_d)
        K.lmat_apply(a, L, post=self.beam_l_d, out=out, accumulate=True)
        return out

    def wiener_start(self, d):
        """CG starting point x0 = C_raw B C_tot^-1 Y^dag d (weighting.py:174), d (nf, Npix) on the device."""
        a = self.base.to_lm(d)
        t = K.lmat_apply(a, self.base.lmax, mat=self.invCtot_d, post=self.beam_l_d)
        return K.lmat_apply(t, self.base.lmax, mat=self.Cl_l_d)

    def applySinv_optimal(self, input_map, preconditioner='pseudo_inverse', nstep=25, thresh=1e-4, verb=False,
                          input_type='map', lmax=None):
        """Optimal S^-1 by CG from the Wiener start (weighting.py:148-184)."""
        assert input_type == 'map', "Optimal weighting should be applied only to pixel-space data!"
        if not hasattr(self, 'noise_cov'):
            raise Exception("The CGD solver requires a pixel-space noise covariance")
        if lmax is None:
            lmax = self.base.lmax
        npy = not isinstance(input_map, torch.Tensor)
        d = input_map if not npy else torch.from_numpy(np.ascontiguousarray(input_map))
        d = d.to(self.base.device, self.base.rdtype).reshape(self.nf, -1).contiguous()
        x0 = self.wiener_start(d)
        if preconditioner == 'learned':
            # flexible CG with the learned preconditioner (no counterpart in the reference)
            from .learned import FlexibleCGD
            if getattr(self, 'learned', None) is None:
                raise Exception("preconditioner='learned' needs a trained LearnedPreconditioner in self.learned")
            self.learned.enabled = True
            cgd = FlexibleCGD(self, d, self.learned, x0=x0, verb=verb)
        else:
            cgd = CGD(self, d, x0=x0, preconditioner=preconditioner, verb=verb)
        cgd.run(nstep=nstep, thresh=thresh)
        self.last_cg = cgd
        out = cgd.get_ivar_lm()
        if lmax != self.base.lmax:
            out = truncate_lm(out, self.base.lmax, lmax)
        return self.base._ret(out, npy)


    # ------------------------------------------------ batched optimal S^-1
    def applySinv_optimal_batch(self, input_maps, preconditioner='pseudo_inverse', nstep=25, thresh=1e-4,
                                verb=False, lmax=None):
        """applySinv_optimal of every map set of a stack (B, nf, Npix) -> (B, nf, nlm(lmax)).

        The B systems run the reference's CG (weighting.py:148-184, 326-351) in lock step with
        their own step lengths; each stops at the iteration where its own error falls below
        thresh, as a separate call would, while the transforms of all B run batched."""
        if not hasattr(self, 'noise_cov'):
            raise Exception("The CGD solver requires a pixel-space noise covariance")
        if preconditioner not in ('harmonic', 'pseudo_inverse'):
            raise Exception("batched CG supports the 'harmonic' and 'pseudo_inverse' preconditioners")
        if lmax is None:
            lmax = self.base.lmax
        npy = not isinstance(input_maps, torch.Tensor)
        d = input_maps if not npy else torch.from_numpy(np.ascontiguousarray(input_maps))
        d = d.to(self.base.device, self.base.rdtype).reshape(d.shape[0], self.nf, -1).contiguous()
        cgd = BatchCGD(self, d, x0=self.wiener_start_batch(d), preconditioner=preconditioner, verb=verb)
        cgd.run(nstep=nstep, thresh=thresh)
        self.last_cg = cgd
        out = cgd.get_ivar_lm()
        if lmax != self.base.lmax:
            out = truncate_lm(out, self.base.lmax, lmax)
        return self.base._ret(out, npy)

    def wiener_start_batch(self, d):
        a = self.base.to_lm_batch(d)
        t = K.lmat_apply_batch(a, self.base.lmax, mat=self.invCtot_d, post=self.beam_l_d)
        return K.lmat_apply_batch(t, self.base.lmax, mat=self.Cl_l_d)

    def apply_A_batch(self, x):
        """apply_A on every system of a batch, x (B, nf, nlm(base_lmax))."""
        L = self.base.lmax
        bx = K.lmat_apply_batch(x, L, post=self.beam_l_d)
        m = self.base.to_map_batch(bx)
        m = K.multiplyRR_batch(m, self.wpix_d)
        a = self.base.to_lm_batch(m)
        out = K.lmat_apply_batch(x, L, mat=self.inv_Cl_l_d)
        K.lmat_apply_batch(a, L, post=self.beam_l_d, out=out, accumulate=True)
        return out


_trunc_idx = {}


def truncate_lm(a, lmax_in, lmax_out):
    """Restrict m-major alm (..., nlm(lmax_in)) to l <= lmax_out (the lfilt selections of the reference)."""
    key = (lmax_in, lmax_out, a.device)
    idx = _trunc_idx.get(key)
    if idx is None:
        idx = [np.arange(m * (2 * lmax_in + 1 - m) // 2 + m, m * (2 * lmax_in + 1 - m) // 2 + lmax_out + 1)
               for m in range(lmax_out + 1)]
        idx = torch.from_numpy(np.concatenate(idx)).to(a.device)
        _trunc_idx[key] = idx
    return a.index_select(-1, idx).contiguous()


class CGD:
    """Conjugate-gradient solver of A x = b (weighting.py:217-367) on the device."""

    def __init__(self, weightings, d, x0=[], test_b=[], preconditioner='pseudo_inverse', verb=False):
        self.preconditioner = preconditioner
        self.base = weightings.base
        self.weightings = weightings
        self.verb = verb
        self.L = self.base.lmax
        W = weightings
        if len(test_b) == 0:
            b0 = self._get_b(d)
        else:
            print("Running in test mode!")
            assert len(d) == 0, "Cannot specify both d and test_b!"
            b0 = W._dev_vec(test_b).clone()
        if len(x0) != 0:
            self.x0 = W._dev_vec(x0)
            b = b0 - self._apply_A(self.x0)
        else:
            self.x0 = torch.zeros_like(b0)
            b = b0
        self.b0 = b0
        self.x = torch.zeros_like(b)
        self.r = b.clone()
        z = self._apply_M(self.r)
        self.rz = self.dot(self.r, z)
        self.rz0 = self.dot(b0, self._apply_M(b0))
        self.p = z
        self.z = z
        self.i = 0
        self.bnorm = float(torch.sqrt(self.dot(b0, b0)))
        self.err = float(self.rz / self.rz0)
        self.history = []
        self.callback = None
        self.keep_Ap = False

    def dot(self, a, b):
        """Re sum a conj(b) (1 + [m > 0]) in FP64 (weighting.py:255); device scalar."""
        return K.mdot(a, b, self.L)

    # pointing operators (weighting.py:267-294)
    def _get_b(self, d):
        W = self.weightings
        m = K.multiplyRR(d.contiguous(), W.Ninv_d)
        a = self.base.to_lm(m)
        return K.lmat_apply(a, self.L, post=W.beam_l_d)

    def _apply_A(self, x):
        """A x = C^-1 x + B Y^dag W N^-1 W Y B x (weighting.py:297-299)."""
        return self.weightings.apply_A(x)

    def _apply_M(self, r):
        """Preconditioner (weighting.py:301-324): 'harmonic' Sigma r, or 'pseudo_inverse'."""
        W = self.weightings
        if self.preconditioner == 'harmonic':
            return K.lmat_apply(r, self.L, mat=W.Sigma_l_d)
        elif self.preconditioner == 'pseudo_inverse':
            alm = K.lmat_apply(r, self.L, mat=W.Sigma_l_d)
            sig = K.lmat_apply(alm, self.L, mat=W.inv_Cl_l_d)
            t = K.lmat_apply(alm, self.L, post=W.bn_d)
            m = self.base.to_map(t)
            m = K.multiplyRR(m, W.noise_cov_d)
            a = self.base.to_lm(m)
            K.lmat_apply(a, self.L, post=W.bnA_d, out=sig, accumulate=True)
            return K.lmat_apply(sig, self.L, mat=W.Sigma_l_d)
        else:
            raise Exception("Unknown preconditioner %s specified!" % self.preconditioner)

    def step(self):
        """One CG iteration (weighting.py:326-340), Fletcher-Reeves beta."""
        Ap = self._apply_A(self.p)
        pAp = self.dot(self.p, Ap)
        if self.keep_Ap:
            # exact (A p, p) pair of this iteration, for training a learned preconditioner
            self.last_p, self.last_Ap = self.p.clone(), Ap
        K.cg_update_xr(self.rz, pAp, self.p, Ap, self.x, self.r)
        del Ap
        z = self._apply_M(self.r)
        next_rz = self.dot(self.r, z)
        K.cg_update_p(next_rz, self.rz, z, self.p)
        self.rz = next_rz
        self.z = z
        self.i += 1
        # the statistic and |r| reach the host in one transfer (one synchronisation per iteration)
        err, rr = torch.stack([self.rz / self.rz0, self.dot(self.r, self.r)]).cpu().tolist()
        self.err = err
        self.history.append((self.i, self.err, float(np.sqrt(rr)) / self.bnorm))
        if self.callback is not None:
            self.callback(self)

    def run(self, nstep=10, thresh=1e-6):
        """Iterate until rz/rz0 < thresh (weighting.py:342-351)."""
        i = -1
        for i in range(nstep):
            self.step()
            if self.verb:
                print("CGD step %d: error = %.2e" % (self.i, self.err))
            if self.err < thresh:
                if self.verb:
                    print("CGD terminating at step %d" % self.i)
                break
        if i == nstep - 1 and self.err > thresh:
            print("CGD did not converge to %.2e threshold after %d steps. Final error = %.2e" % (thresh, nstep,
                                                                                               self.err))

    def get_wiener_lm(self):
        return self.x + self.x0

    def get_wiener_map(self):
        return self.base.to_map(self.x + self.x0)

    def get_ivar_lm(self):
        return K.lmat_apply((self.x + self.x0).contiguous(), self.L, mat=self.weightings.inv_Cl_l_d)

    def get_ivar_map(self):
        return self.base.to_map(self.get_ivar_lm())


class BatchCGD:
    """CGD (weighting.py:217-367) for B systems at once: vectors (B, nf, nlm), one step
    length and one error per system, converged systems frozen (their x, r, p unchanged)."""

    def __init__(self, weightings, d, x0=None, preconditioner='pseudo_inverse', verb=False):
        self.preconditioner = preconditioner
        self.base = weightings.base
        self.weightings = weightings
        self.verb = verb
        self.L = self.base.lmax
        W = weightings
        b0 = self._get_b(d)
        if x0 is not None:
            self.x0 = x0
            b = b0 - W.apply_A_batch(x0)
        else:
            self.x0 = torch.zeros_like(b0)
            b = b0
        self.nb = b0.shape[0]
        self.x = torch.zeros_like(b)
        self.r = b.clone()
        z = self._apply_M(self.r)
        self.rz = self.dot(self.r, z)
        self.rz0 = self.dot(b0, self._apply_M(b0))
        self.p = z
        self.active = torch.ones(self.nb, dtype=torch.int32, device=b.device)
        self.i = 0
        self.iterations = np.zeros(self.nb, dtype=np.int64)  # steps each system took
        self.err = (self.rz / self.rz0).cpu().numpy()
        self.history = []

    def dot(self, a, b):
        return K.mdot_batch(a, b, self.L)

    def _get_b(self, d):
        W = self.weightings
        m = K.multiplyRR_batch(d.contiguous(), W.Ninv_d)
        a = self.base.to_lm_batch(m)
        return K.lmat_apply_batch(a, self.L, post=W.beam_l_d)

    def _apply_M(self, r):
        W = self.weightings
        if self.preconditioner == 'harmonic':
            return K.lmat_apply_batch(r, self.L, mat=W.Sigma_l_d)
        alm = K.lmat_apply_batch(r, self.L, mat=W.Sigma_l_d)
        sig = K.lmat_apply_batch(alm, self.L, mat=W.inv_Cl_l_d)
        t = K.lmat_apply_batch(alm, self.L, post=W.bn_d)
        m = self.base.to_map_batch(t)
        m = K.multiplyRR_batch(m, W.noise_cov_d)
        a = self.base.to_lm_batch(m)
        K.lmat_apply_batch(a, self.L, post=W.bnA_d, out=sig, accumulate=True)
        return K.lmat_apply_batch(sig, self.L, mat=W.Sigma_l_d)

    def step(self):
        """One CG iteration of every active system (weighting.py:326-340)."""
        Ap = self.weightings.apply_A_batch(self.p)
        pAp = self.dot(self.p, Ap)
        K.cg_update_xr_batch(self.rz, pAp, self.active, self.p, Ap, self.x, self.r)
        del Ap
        z = self._apply_M(self.r)
        next_rz = self.dot(self.r, z)
        K.cg_update_p_batch(next_rz, self.rz, self.active, z, self.p)
        self.rz = torch.where(self.active.bool(), next_rz, self.rz)
        self.i += 1
        act = self.active.cpu().numpy().astype(bool)
        self.iterations[act] += 1
        self.err = (self.rz / self.rz0).cpu().numpy()
        self.history.append((self.i, self.err.copy()))

    def run(self, nstep=10, thresh=1e-6):
        """Iterate until every system has rz/rz0 < thresh (weighting.py:342-351, per system)."""
        for _ in range(nstep):
            self.step()
            done = self.err < thresh
            if self.verb:
                print("batched CGD step %d: max error = %.2e, %d of %d converged" %
                      (self.i, self.err.max(), int(done.sum()), self.nb))
            if done.any():
                self.active.copy_(torch.from_numpy((~done).astype(np.int32)))
            if done.all():
                break
        left = int((self.err >= thresh).sum())
        if left:
            print("batched CGD: %d of %d systems did not converge to %.2e after %d steps" % (left, self.nb, thresh,
                                                                                          nstep))

    def get_wiener_lm(self):
        return self.x + self.x0

    def get_ivar_lm(self):
        return K.lmat_apply_batch((self.x + self.x0).contiguous(), self.L, mat=self.weightings.inv_Cl_l_d)


# applySinv_ideal issues no host synchronisation, so a numerator that calls it can be captured
# into a HIP graph (TSpecTemplate(capture_graph=True) checks this flag)
Weightings.applySinv_ideal.capturable = True
