"""PolySpec base class with the MI355X ``backend="hip"`` SHT backend.

Mirrors polyspec/base.py (class PolySpec, base.py:9-446): same constructor
arguments, attributes, SHT method names/signatures, shape asserts and SHT
counters, so that PSpecBin / BSpecBin / TSpecTemplate style callers run
unchanged.  The transforms run on the GPU through libpolyspec_hip.so:

  * numpy in -> numpy out (float64 maps, complex128 alm, like the reference);
  * torch device tensors in -> device tensors out (no host round trip).

``precision`` selects the device arithmetic of the bulk transforms: 64 (FP64,
the default, matching the reference) or 32 (FP32 storage/accumulation with
the Legendre recursion in FP64).
"""
import os
import time

import numpy as np
import torch

from . import kernels as K
from ._abi import ShtPlan


def getlm(lmax):
    """(l, m) of every coefficient in healpy's m-major order (healpy.Alm.getlm)."""
    m = np.concatenate([np.full(lmax + 1 - mm, mm) for mm in range(lmax + 1)])
    l = np.concatenate([np.arange(mm, lmax + 1) for mm in range(lmax + 1)])
    return l, m


def nlm_of(lmax):
    return (lmax + 1) * (lmax + 2) // 2


class PolySpec:
    """Base class for PolySpec (polyspec/base.py:9-117) with the HIP SHT backend.

    Inputs: Nside, Cl_tot (dict of fiducial total spectra), beam, pol, backend,
    nthreads -- as in the reference -- plus ``precision`` (64 | 32),
    ``device`` (HIP device index) for the GPU backend and ``autotune``: time the
    launch-shape candidates of each SHT plan on its first transform of a batch width
    and keep the fastest (ShtPlan.autotune; results agree with the built-in shapes to
    rounding).
    """

    def __init__(self, Nside, Cl_tot, beam=[], pol=True, backend='hip', nthreads=0, precision=64, device=None,
                 autotune=True):
        if backend != 'hip':
            raise Exception("This package implements the 'hip' backend only ('healpix' and 'ducc' "
                            "are provided by the reference package)")
        self.Nside = Nside
        self.pol = pol
        self.backend = backend
        self.precision = int(precision)
        self.autotune = bool(autotune)
        assert self.precision in (32, 64), "precision must be 32 or 64"
        if not torch.cuda.is_available():
            raise RuntimeError("backend='hip' needs a visible MI355X (HIP) device")
        self.device = torch.device('cuda', torch.cuda.current_device() if device is None else device)
        self.rdtype = torch.float32 if self.precision == 32 else torch.float64
        self.cdtype = torch.complex64 if self.precision == 32 else torch.complex128

        # Derived parameters (base.py:27-35)
        self.Npix = 12 * Nside ** 2
        self.A_pix = 4. * np.pi / self.Npix
        self.lmax = 3 * self.Nside - 1
        self.l_arr, self.m_arr = getlm(self.lmax)
        self.nthreads = len(os.sched_getaffinity(0)) if nthreads <= 0 else nthreads

        self.indices = {'T': 0, 'E': 1, 'B': 2}
        self.parities = {'T': 1, 'E': 1, 'B': -1}

        # Fiducial spectra (base.py:41-49)
        ls = np.arange(self.lmax + 1)
        if self.pol:
            assert len(Cl_tot.keys()) == 6 or len(Cl_tot.keys()) == 4, \
                "Must specify four or six input spectra: {ClTT, ClTE, ClEE, ClBB} or {ClTT, ClTE, ClTB, ClEE, ClEB, ClBB}"
            if len(Cl_tot.keys()) == 4:
                Cl_tot['TB'] = Cl_tot['EB'] = 0. * Cl_tot['TT']
        for k in Cl_tot.keys():
            assert len(Cl_tot[k]) >= len(ls), "Cl_tot must contain every ell mode from l=0 to l=lmax!"
        self.Cl_tot = Cl_tot.copy()

        # Beam (base.py:51-64)
        if len(beam) == 0:
            self.beam = np.asarray([1. + 0. * self.Cl_tot['TT'] for _ in range(1 + 2 * self.pol)])
        else:
            if self.pol:
                assert len(beam) == 2 or len(beam) == 3, "Beam must contain temperature and polarization or T/E/B components"
                b = [beam[0], beam[1], beam[1]] if len(beam) == 2 else [beam[0], beam[1], beam[2]]
                self.beam = np.asarray([np.array(x, copy=True) for x in b])
            else:
                assert (len(beam) == 1 or len(beam) == len(Cl_tot['TT'])), "Beam must contain the same ells as Cl_tot"
                self.beam = np.asarray(beam).copy().reshape(1, -1)
        for i in range(len(self.beam)):
            assert len(self.beam[i]) >= len(ls), "Beam must contain every ell mode from l=0 to l=lmax!"

        self.m_weight = (1. + 1. * (self.m_arr > 0.))

        # C^-1 per l, zero where the covariance is not positive (base.py:91-106)
        if pol:
            C = np.zeros((len(ls), 3, 3))
            for a, ka in enumerate('TEB'):
                for b, kb in enumerate('TEB'):
                    key = ka + kb if ka + kb in self.Cl_tot else kb + ka
                    C[:, a, b] = self.Cl_tot[key][:len(ls)]
            good = np.linalg.det(C) > 0
            self.inv_Cl_tot_mat = np.zeros((3, 3, len(ls)))
            self.inv_Cl_tot_mat[:, :, good] = np.transpose(np.linalg.inv(C[good]), (1, 2, 0))
        else:
            good = self.Cl_tot['TT'][:len(ls)] > 0
            self.inv_Cl_tot_mat = np.zeros((1, 1, len(ls)))
            self.inv_Cl_tot_mat[0, 0, good] = 1. / self.Cl_tot['TT'][:len(ls)][good]
        self.inv_Cl_tot_lm_mat = np.ascontiguousarray(self.inv_Cl_tot_mat[:, :, self.l_arr])
        self.beam_lm = np.ascontiguousarray(self.beam[:, self.l_arr])

        self.n_SHTs_forward = 0
        self.n_SHTs_reverse = 0
        self.time_sht = 0.
        self._plans = {}
        # draw Gaussian random fields with the device generator (fast, large maps)
        # instead of numpy on the host
        self.grf_device = False
        self._grf_cache = {}  # synalm's per-l factors and index arrays by (lmax, spectra)

    # ------------------------------------------------------------ plumbing
    def timer_func(func):
        def wrap_func(self, *args, **kwargs):
            t1 = time.time()
            result = func(self, *args, **kwargs)
            self.time_sht += time.time() - t1
            return result
        wrap_func.__name__ = func.__name__
        wrap_func.__doc__ = func.__doc__
        return wrap_func

    def plan(self, lmax, spin):
        key = (int(lmax), int(spin))
        p = self._plans.get(key)
        if p is None:
            with torch.cuda.device(self.device):
                p = ShtPlan(self.Nside, lmax, spin, self.precision, autotune=self.autotune)
            self._plans[key] = p
        return p

    def _to_dev(self, x, complex_):
        dt = self.cdtype if complex_ else self.rdtype
        if isinstance(x, torch.Tensor):
            return x.to(self.device, dt).contiguous()
        return torch.from_numpy(np.ascontiguousarray(x)).to(self.device, dt)

    @staticmethod
    def _ret(t, like_numpy):
        if not like_numpy:
            return t
        t = t.cpu()
        if t.is_complex():
            return t.to(torch.complex128).numpy()
        return t.to(torch.float64).numpy()

    # ------------------------------------------------------ SHT interface
    @timer_func
    def to_lm(self, input_map, lmax=None):
        """Map -> harmonic space (TQU -> TEB for three fields), base.py:130-151.

        adjoint synthesis times pi/(3 Nside^2) = 4 pi/Npix (no iteration)."""
        self.n_SHTs_forward += 1 + (len(input_map) == 3)
        assert len(input_map) in [1, 3], "Wrong input shape supplied!"
        if lmax is None:
            lmax = self.lmax
        npy = not isinstance(input_map, torch.Tensor)
        m = self._to_dev(input_map, False).reshape(len(input_map), self.Npix)
        sc = np.pi / (3. * self.Nside ** 2)
        p0 = self.plan(lmax, 0)
        out = torch.empty((len(input_map), p0.nlm), dtype=self.cdtype, device=self.device)
        if len(input_map) == 3:
            # T (spin 0) on a side stream with its own workspace, Q/U (spin 2) on the current
            # stream: the two single-map adjoints are latency-bound and overlap.  Under graph
            # capture the fork and join become graph dependencies.
            p2 = self.plan(lmax, 2)
            main = torch.cuda.current_stream(self.device)
            side = self._side_stream()
            ready = torch.cuda.Event()
            ready.record(main)
            side.wait_event(ready)
            ws = self._side_workspace(p0.work_bytes(1, 1))
            if not torch.cuda.is_current_stream_capturing():
                m.record_stream(side)  # used on the side stream: not reused before its work ends
                out.record_stream(side)
            with torch.cuda.stream(side):
                p0.adjoint(m[0:1].reshape(1, 1, -1), sc, out=out[0:1].view(1, 1, -1), ws=ws)
                done = torch.cuda.Event()
                done.record(side)
            p2.adjoint(m[1:3].reshape(1, 2, -1), sc, out=out[1:3].view(1, 2, -1))
            main.wait_event(done)
        else:
            p0.adjoint(m[0:1].reshape(1, 1, -1), sc, out=out[0:1].view(1, 1, -1))
        return self._ret(out, npy)

    def _side_stream(self):
        """A second stream (default priority) for transforms that overlap with the current one."""
        if getattr(self, '_side', None) is None:
            self._side = torch.cuda.Stream(device=self.device)
        return self._side

    def _side_workspace(self, nbytes):
        """Workspace of the side stream (the shared one belongs to the current stream)."""
        buf = getattr(self, '_side_ws', None)
        if buf is None or buf.numel() < nbytes:
            self._side_ws = None
            # allocated on the side stream so the caching allocator orders its reuse there
            with torch.cuda.stream(self._side_stream()):
                self._side_ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=self.device)
        return self._side_ws

    @timer_func
    def to_map(self, input_lm, lmax=None):
        """Harmonic -> map space (TEB -> TQU for three fields), base.py:153-176."""
        self.n_SHTs_reverse += 1 + (len(input_lm) == 3)
        assert len(input_lm) in [1, 3], "Wrong input shape supplied!"
        if lmax is None:
            lmax = self.lmax
        assert len(input_lm[0]) == (lmax + 1) * (lmax + 2) // 2, "Wrong number of alm coefficients!"
        npy = not isinstance(input_lm, torch.Tensor)
        a = self._to_dev(input_lm, True)
        out = torch.empty((len(input_lm), self.Npix), dtype=self.rdtype, device=self.device)
        p0 = self.plan(lmax, 0)
        if len(input_lm) == 3:
            # T on a side stream with its own workspace, E/B on the current stream (as in to_lm)
            p2 = self.plan(lmax, 2)
            main = torch.cuda.current_stream(self.device)
            side = self._side_stream()
            ready = torch.cuda.Event()
            ready.record(main)
            side.wait_event(ready)
            ws = self._side_workspace(p0.work_bytes(1, 0))
            if not torch.cuda.is_current_stream_capturing():
                a.record_stream(side)
                out.record_stream(side)
            with torch.cuda.stream(side):
                p0.synthesis(a[0:1].reshape(1, 1, -1), out=out[0:1].view(1, 1, -1), ws=ws)
                done = torch.cuda.Event()
                done.record(side)
            p2.synthesis(a[1:3].reshape(1, 2, -1), out=out[1:3].view(1, 2, -1))
            main.wait_event(done)
        else:
            p0.synthesis(a[0:1].reshape(1, 1, -1), out=out[0:1].view(1, 1, -1))
        return self._ret(out, npy)

    def to_lm_batch(self, maps, lmax=None, tune=True):
        """to_lm of a stack of map sets, (B, nf, Npix) -> (B, nf, nlm), nf = 1 or 3: one batched
        adjoint per spin over all B sets (the reference transforms one set per call).  tune=False
        keeps the built-in launch shapes for a batch width the plans have not tuned."""
        B, nf = maps.shape[0], maps.shape[1]
        assert nf in (1, 3), "Wrong input shape supplied!"
        lmax = self.lmax if lmax is None else lmax
        self.n_SHTs_forward += B * (1 + (nf == 3))
        m = self._to_dev(maps, False).reshape(B, nf, self.Npix)
        sc = np.pi / (3. * self.Nside ** 2)
        p0 = self.plan(lmax, 0)
        plans = [p0] if nf == 1 else [p0, self.plan(lmax, 2)]
        if not tune:
            for p in plans:
                p.skip_tuning(B)
        if nf == 1:
            return p0.adjoint(m, sc)
        out = torch.empty((B, 3, p0.nlm), dtype=self.cdtype, device=self.device)
        out[:, 0:1] = p0.adjoint(m[:, 0:1].contiguous(), sc)
        out[:, 1:3] = plans[1].adjoint(m[:, 1:3].contiguous(), sc)
        return out

    def to_map_batch(self, alms, lmax=None):
        """to_map of a stack of harmonic sets, (B, nf, nlm) -> (B, nf, Npix), nf = 1 or 3."""
        B, nf = alms.shape[0], alms.shape[1]
        assert nf in (1, 3), "Wrong input shape supplied!"
        lmax = self.lmax if lmax is None else lmax
        self.n_SHTs_reverse += B * (1 + (nf == 3))
        a = self._to_dev(alms, True)
        p0 = self.plan(lmax, 0)
        if nf == 1:
            return p0.synthesis(a)
        out = torch.empty((B, 3, self.Npix), dtype=self.rdtype, device=self.device)
        out[:, 0:1] = p0.synthesis(a[:, 0:1].contiguous())
        out[:, 1:3] = self.plan(lmax, 2).synthesis(a[:, 1:3].contiguous())
        return out

    @timer_func
    def to_lm_vec(self, input_vec, spin=0, lmax=None):
        """Batched map -> harmonic transform, base.py:178-209.

        spin 0: (B, Npix) -> (B, nlm).  spin s > 0: input [(+s) maps, (-s) maps],
        each (B, Npix) complex, output [(+s) alm, (-s) alm] each (B, nlm)."""
        self.n_SHTs_forward += len(input_vec)
        if lmax is None:
            lmax = self.lmax
        sc = np.pi / (3. * self.Nside ** 2)
        npy = not isinstance(input_vec, torch.Tensor)
        if spin == 0:
            m = self._to_dev(input_vec, False)
            B = m.shape[0]
            out = self.plan(lmax, 0).adjoint(m.reshape(B, 1, -1), sc)
            return self._ret(out.view(B, -1), npy)
        assert spin > 0, "Spin must be positive!"
        assert len(input_vec) == 2, "Must input both +s and -s maps"
        P = input_vec[0] if isinstance(input_vec[0], torch.Tensor) else torch.from_numpy(np.asarray(input_vec[0]))
        M = input_vec[1] if isinstance(input_vec[1], torch.Tensor) else torch.from_numpy(np.asarray(input_vec[1]))
        P = P.to(self.device, self.cdtype)
        M = M.to(self.device, self.cdtype)
        P = P.reshape(-1, self.Npix)
        M = M.reshape(-1, self.Npix)
        # (ideal_fisher.to_real_imag, ideal_fisher.pyx:308-322): Re (P+M)/2, Im (P-M)/2
        ri = torch.stack([(P + M).real / 2., (P - M).imag / 2.], dim=1).contiguous()
        GC = self.plan(lmax, spin).adjoint(ri, sc)
        # (to_plus_minus_complex, ideal_fisher.pyx:254-271)
        plus = -(GC[:, 0] + 1j * GC[:, 1])
        minus = -(GC[:, 0] - 1j * GC[:, 1]) if spin % 2 == 0 else (GC[:, 0] - 1j * GC[:, 1])
        out = torch.stack([plus, minus])
        return self._ret(out, npy)

    @timer_func
    def to_map_vec(self, input_lm_vec, output_spin=0, lmax=None):
        """Batched harmonic -> map transform, base.py:211-232."""
        self.n_SHTs_reverse += len(input_lm_vec)
        if lmax is None:
            lmax = self.lmax
        assert len(input_lm_vec[0]) == (lmax + 1) * (lmax + 2) // 2, "Wrong number of alm coefficients!"
        npy = not isinstance(input_lm_vec, torch.Tensor)
        a = self._to_dev(input_lm_vec, True)
        B = a.shape[0]
        if output_spin == 0:
            out = self.plan(lmax, 0).synthesis(a.reshape(B, 1, -1))
            return self._ret(out.view(B, -1), npy)
        GC = torch.stack([-a, torch.zeros_like(a)], dim=1).contiguous()
        RI = self.plan(lmax, output_spin).synthesis(GC)
        # (ideal_fisher.to_plus_minus, ideal_fisher.pyx:237-251)
        out = torch.stack([RI[:, 0] + 1j * RI[:, 1], RI[:, 0] - 1j * RI[:, 1]])
        return self._ret(out, npy)

    @timer_func
    def to_lm_spin(self, input_map_plus, input_map_minus, spin, lmax=None):
        """(+-s) maps -> (+-s) harmonics with HEALPix conventions, base.py:234-258."""
        assert spin > 0, "Spin must be positive!"
        assert type(spin) == int, "Spin must be an integer!"
        self.n_SHTs_forward += 1
        if lmax is None:
            lmax = self.lmax
        npy = not isinstance(input_map_plus, torch.Tensor)
        P = self._to_dev(input_map_plus, True).reshape(self.Npix)
        M = self._to_dev(input_map_minus, True).reshape(self.Npix)
        ri = torch.stack([((P + M) / 2.).real, ((P - M) / 2j).real]).reshape(1, 2, -1).contiguous()
        GC = self.plan(lmax, spin).adjoint(ri, np.pi / (3. * self.Nside ** 2))[0]
        lm_plus = -(GC[0] + 1j * GC[1])
        lm_minus = -1 * (-1) ** spin * (GC[0] - 1j * GC[1])
        return self._ret(torch.stack([lm_plus, lm_minus]), npy)

    @timer_func
    def to_map_spin(self, _input_lm_plus, _input_lm_minus, spin, lmax=None):
        """(+-s) harmonics -> (+-s) maps with HEALPix conventions, base.py:260-288."""
        assert spin > 0, "Spin must be positive!"
        assert type(spin) == int, "Spin must be an integer!"
        self.n_SHTs_reverse += 1
        if lmax is None:
            lmax = self.lmax
        assert len(_input_lm_plus) == (lmax + 1) * (lmax + 2) // 2, "Wrong number of alm coefficients!"
        npy = not isinstance(_input_lm_plus, torch.Tensor)
        lp = self._to_dev(_input_lm_plus, True)
        lm = self._to_dev(_input_lm_minus, True)
        sg = (-1) ** spin
        G = -(lp + sg * lm) / 2.
        C = -(lp - sg * lm) / 2j
        RI = self.plan(lmax, spin).synthesis(torch.stack([G, C]).reshape(1, 2, -1).contiguous())[0]
        map_plus = RI[0] + 1j * RI[1]
        map_minus = RI[0] - 1j * RI[1]
        return self._ret(map_plus, npy), self._ret(map_minus, npy)

    def compute_spin_transform_map(self, a_lms, spin, lmax=None):
        """sum_lm sY_lm a_lm for scalar a_lm (base.py:290-297)."""
        if spin > 0:
            return [self.to_map_spin(a_lm, (-1.) ** spin * a_lm, int(abs(spin)), lmax=lmax)[0] for a_lm in a_lms]
        if spin < 0:
            return [self.to_map_spin((-1.) ** spin * a_lm, a_lm, int(abs(spin)), lmax=lmax)[1] for a_lm in a_lms]

    def safe_divide(self, x, y):
        """Divide maps without zero errors (base.py:299-303)."""
        out = np.zeros_like(x)
        out[y != 0] = x[y != 0] / y[y != 0]
        return out

    # ---------------------------------------------------------- utilities
    def _grf_factors(self, Cl_input, lmax):
        """Per-l factors L_l with L_l L_l^T = C_l for the spectra Cl_input (an eigen-decomposition
        square root, also valid for singular C_l), with the l and m of every (l, m): numpy arrays,
        and with grf_device the factors, the l index and the m = 0 mask on the device.  Cached by
        lmax and the spectra's values, so that a run of draws does no per-draw host work (a
        per-l decomposition loop and host-to-device copies of the index arrays stalled the device
        queue between draws)."""
        nf = 3 if self.pol else 1
        keys = [['TT', 'TE', 'TB'], ['TE', 'EE', 'EB'], ['TB', 'EB', 'BB']]
        C = np.zeros((lmax + 1, nf, nf))
        for a in range(nf):
            for b in range(nf):
                k = keys[a][b]
                if k in Cl_input:
                    C[:, a, b] = np.asarray(Cl_input[k])[:lmax + 1]
                elif k[::-1] in Cl_input:
                    C[:, a, b] = np.asarray(Cl_input[k[::-1]])[:lmax + 1]  # TB/EB may be absent: zero
        key = (lmax, C.tobytes())
        ent = self._grf_cache.get(key)
        if ent is None:
            w, v = np.linalg.eigh(C)  # every l at once
            ls, ms = getlm(lmax)
            ent = {'Lc': v * np.sqrt(np.clip(w, 0, None))[:, None, :], 'ls': ls, 'ms': ms}
            if len(self._grf_cache) >= 4:
                self._grf_cache.clear()
            self._grf_cache[key] = ent
        if self.grf_device and 'Lt' not in ent:
            ent['Lt'] = torch.from_numpy(ent['Lc']).to(self.device)
            ent['li'] = torch.from_numpy(ent['ls']).to(self.device)
            ent['m0'] = torch.from_numpy(ent['ms'] == 0).to(self.device)
        return ent

    def synalm(self, seed, Cl_input, lmax=None):
        """Seeded Gaussian a_lm with covariance C_l (the role of healpy.synalm at
        base.py:332-334; this generator's random stream is its own).

        Per l, a square-root factor of the (1x1 or 3x3) C_l draws correlated
        T[,E,B]; m = 0 coefficients are real with variance C_l, m > 0 have
        independent real/imaginary parts of variance C_l/2."""
        lmax = self.lmax if lmax is None else lmax
        nf = 3 if self.pol else 1
        f = self._grf_factors(Cl_input, lmax)
        ls, ms = f['ls'], f['ms']
        if self.grf_device:
            g = torch.Generator(device=self.device)
            g.manual_seed(int(seed) if seed is not None else 0)
            zr = torch.randn((nf, len(ls)), dtype=torch.float64, device=self.device, generator=g)
            zi = torch.randn((nf, len(ls)), dtype=torch.float64, device=self.device, generator=g)
            m0 = f['m0']
            zr = torch.where(m0, zr * np.sqrt(2.), zr)
            zi = torch.where(m0, torch.zeros_like(zi), zi)
            z = torch.complex(zr, zi) / np.sqrt(2.)
            # out[a] = sum_b Lc[l, a, b] z[b]: elementwise over (l, m) (a batched 3 x 3 product per
            # (l, m) as a library GEMM took 16 ms per Nside-1024 draw)
            Lt, li = f['Lt'], f['li']
            out = torch.empty((nf, len(ls)), dtype=torch.complex128, device=self.device)
            for a in range(nf):
                acc = z[0] * Lt[:, a, 0][li]
                for b in range(1, nf):
                    acc = acc + z[b] * Lt[:, a, b][li]
                out[a] = acc
            return out
        rng = np.random.default_rng(seed)
        z = rng.standard_normal((nf, len(ls))) + 1j * rng.standard_normal((nf, len(ls)))
        z[:, ms == 0] = z[:, ms == 0].real * np.sqrt(2.)
        z /= np.sqrt(2.)
        return np.einsum('lab,bl->al', f['Lc'][ls], z)

    def generate_data(self, seed=None, Cl_input={}, output_type='map', deconvolve_beam=False, b_input=None,
                      add_B=False, remove_mean=True, sum_ells='even', lmax=None):
        """Full-sky Gaussian random field with spectra Cl_input (base.py:305-343).

        The bispectrum-adding branch (add_B) of the reference is not part of
        this backend (it is not on the gNL/tauNL-loc hot path)."""
        assert output_type in ['harmonic', 'map'], "Valid output types are 'harmonic' and 'map' only!"
        if add_B:
            raise NotImplementedError("generate_data(add_B=True) is not provided by the hip backend")
        if lmax is None:
            lmax = self.lmax
        if len(Cl_input.keys()) == 0:
            Cl_input = self.Cl_tot
        if self.pol:
            assert len(Cl_input.keys()) == 6 or len(Cl_input.keys()) == 4, \
                "Need to specify {ClTT, ClEE, ClBB, ClTE} and optionally {ClTB, ClEB}."
            if len(Cl_input.keys()) == 4:
                Cl_input['TB'] = Cl_input['EB'] = 0. * Cl_input['TT']
        initial_lm = self.synalm(seed, Cl_input, self.lmax)
        sel = self.l_arr <= lmax if lmax < self.lmax else None  # None: every (l, m) kept
        if sel is None:
            pass
        elif isinstance(initial_lm, torch.Tensor):
            initial_lm = initial_lm[:, torch.from_numpy(sel).to(self.device)].contiguous()
        else:
            initial_lm = initial_lm[:, sel]
        if deconvolve_beam:
            this_beam = self.beam_lm if sel is None else self.beam_lm[:, sel]
            inv = np.where(this_beam != 0, 1. / np.where(this_beam != 0, this_beam, 1.), 0.)
            if isinstance(initial_lm, torch.Tensor):
                initial_lm = initial_lm * torch.from_numpy(inv).to(self.device)
            else:
                initial_lm = initial_lm * inv
        if output_type == 'map':
            return self.to_map(initial_lm, lmax=lmax)
        return initial_lm

    def applyAinv(self, input_map, input_type='map', lmax=-1):
        """Exact inverse weighting B C_tot^-1 B in harmonic space (base.py:424-446)."""
        if lmax == -1:
            lmax = self.lmax
        assert input_type in ['harmonic', 'map'], "Valid input types are 'harmonic' and 'map' only!"
        npy = not isinstance(input_map, torch.Tensor)
        if input_type == 'map':
            input_map_lm = self.to_lm(input_map, lmax=lmax)
        else:
            input_map_lm = input_map
        a = self._to_dev(input_map_lm, True)
        assert a.dim() == 2 and a.shape[1] == (lmax + 1) * (lmax + 2) // 2, "Wrong number of alm coefficients!"
        nf = a.shape[0]
        # B C_tot^-1 B per (l, m) as one device pass of the per-l 3x3 product
        out = K.lmat_apply(a, lmax, mat=np.ascontiguousarray(self.inv_Cl_tot_mat[:nf, :nf, :lmax + 1]),
                           pre=np.ascontiguousarray(self.beam[:nf, :lmax + 1]),
                           post=np.ascontiguousarray(self.beam[:nf, :lmax + 1]))
        return self._ret(out, npy)
