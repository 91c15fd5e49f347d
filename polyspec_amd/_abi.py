"""ctypes bindings to libpolyspec_hip.so (declared in include/polyspec_hip.h).

The product path has no CPU fallback: if the HIP library is missing, importing
the bindings raises.  torch tensors are used only as device containers; every
compute call goes through the C ABI on torch's current HIP stream.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpolyspec_hip.so")

_P, _I, _LL, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double

# name -> (restype, argtypes); every entry point of include/polyspec_hip.h
SIGNATURES = {
    "ps_last_error": (ctypes.c_char_p, []),
    "ps_version": (_I, []),
    "ps_init": (_I, [_I]),
    "ps_sht_plan_create": (_I, [_I, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "ps_sht_plan_destroy": (_I, [_P]),
    "ps_sht_plan_info": (_I, [_P, ctypes.POINTER(_LL), ctypes.POINTER(_LL), ctypes.POINTER(_I)]),
    "ps_sht_workspace_bytes": (_LL, [_P, _I, _I]),
    "ps_sht_lmajor_multiple": (_I, [_P]),
    "ps_sht_synthesis": (_I, [_P, _I, _P, _P, _P, _LL, _P]),
    "ps_sht_synthesis_grad": (_I, [_P, _I, _P, _P, _P, _LL, _P]),
    "ps_sht_synthesis_lmajor": (_I, [_P, _I, _P, _I, _P, _P, _LL, _P]),
    "ps_sht_adjoint": (_I, [_P, _I, _P, _P, _D, _P, _LL, _P]),
    "ps_sht_phase_bytes": (_I, [_P, _I, ctypes.POINTER(_LL)]),
    "ps_sht_shape_candidates": (_I, [_P, _I, _I, ctypes.POINTER(_I), _I]),
    "ps_sht_get_shapes": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "ps_sht_set_shapes": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "ps_sht_autotune_work_bytes": (_LL, [_P, _I]),
    "ps_sht_autotune": (_I, [_P, _I, _P, _LL, _I, ctypes.POINTER(_D), _P]),
    "ps_sht_legendre_synthesis_lmajor": (_I, [_P, _I, _P, _I, _P, _P]),
    "ps_sht_ring_synthesis": (_I, [_P, _I, _P, _P, _P]),
    "ps_reduce_scratch_bytes": (_LL, []),
    "ps_weight_alm_lmajor": (_I, [_I, _P, _I, _I, _P, _I, _I, _I, _D, _I, _I, _P, _P]),
    "ps_weight_alm_pq_lmajor": (_I, [_I, _P, _P, _I, _I, _P, _I, _I, _I, _D, _D, _I, _I, _P, _P]),
    "ps_weight_alm_pq_interleaved_lmajor": (_I, [_I, _P, _P, _I, _I, _P, _I, _I, _I, _D, _D, _I, _I, _P, _P]),
    "ps_sht_ring_gnl_work_bytes": (_LL, [_P, _I]),
    "ps_sht_ring_gnl": (_I, [_P, _I, _P, _P, _P, _P, _P, _D, _D, _P, _LL, _P, _I, _P]),
    "ps_gnl_loc_fused_sum": (_I, [_I, _P, _P, _P, _P, _P, _I, _LL, _D, _D, _P, _P, _I, _P]),
    "ps_gnl_loc_disc_sum": (_I, [_I, _P, _P, _P, _P, _P, _I, _LL, _P, _P, _I, _P]),
    "ps_mc_accumulate": (_I, [_I, _P, _P, _I, _LL, _P, _P, _P]),
    "ps_pixel_products": (_I, [_I, _I, _P, _P, _P, _P, _LL, _D, _P, _P]),
    "ps_radial_sum": (_I, [_I, _P, _I, _I, _I, _P, _I, _I, _P, _D, _I, _P, _P]),
    "ps_radial_scatter": (_I, [_I, _P, _I, _I, _I, _P, _I, _I, _P, _D, _I, _P, _P]),
    "ps_tau_sum_n0": (_I, [_I, _P, _P, _I, _I, _I, _I, _I, _P, _P, _I, _P, _P, _I, _P]),
    "ps_tau_sum_lmajor": (_I, [_I, _P, _P, _LL, _I, _I, _I, _I, _I, _P, _P, _I, _P, _P, _I, _P]),
    "ps_tau_sum_lmajor_noalloc": (_I, [_I, _P, _P, _LL, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P]),
    "ps_tau_sum_lmajor_shift": (_I, [_I, _P, _P, _LL, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _I, _P, _P, _I, _P]),
    "ps_sht_alm_to_lmajor": (_I, [_P, _I, _P, _I, _P, _P]),
    "ps_outer_product": (_I, [_I, _P, _P, _I, _I, _LL, _P, _P, _I, _P]),
    "ps_convolve_F_lmajor": (_I, [_I, _P, _I, _I, _I, _I, _P, _P, _I, _P, _P]),
    "ps_lmat_apply": (_I, [_I, _I, _I, _P, _P, _P, _P, _LL, _P, _LL, _D, _I, _P]),
    "ps_mdot_scratch_bytes": (_LL, []),
    "ps_mdot": (_I, [_I, _I, _I, _P, _P, _P, _P, _P]),
    "ps_cg_update_xr": (_I, [_I, _LL, _P, _P, _P, _P, _P, _P, _P]),
    "ps_cg_update_p": (_I, [_I, _LL, _P, _P, _P, _P, _P]),
    "ps_pixel_products_batch": (_I, [_I, _I, _I, _P, _P, _P, _P, _LL, _D, _P, _P]),
    "ps_lmat_apply_batch": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _D, _I, _P]),
    "ps_mdot_batch_scratch_bytes": (_LL, [_I]),
    "ps_mdot_batch": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "ps_cg_update_xr_batch": (_I, [_I, _I, _LL, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ps_cg_update_p_batch": (_I, [_I, _I, _LL, _P, _P, _P, _P, _P, _P]),
    "ps_nn_im2col9": (_I, [_P, _P, _I, _I, _I, _P, _P]),
    "ps_nn_col2im9": (_I, [_P, _P, _I, _I, _I, _I, _P, _P]),
    "ps_nn_gemm": (_I, [_I, _I, _I, _I, _I, _P, _LL, _P, _LL, _P, _LL, _P, _I, _I, _P, _P]),
    "ps_nn_relu_backward": (_I, [_LL, _P, _P, _P]),
    "ps_nn_colsum_work_floats": (_LL, [_I, _I]),
    "ps_nn_colsum": (_I, [_I, _I, _P, _P, _P, _P]),
    "ps_nn_adam": (_I, [_LL, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _I, _P]),
    "ps_nn_mse": (_I, [_LL, _I, _P, _P, _P, _P, _P, _P, _P]),
    "ps_nn_pixel_scale": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "ps_nn_conv9_bf16": (_I, [_P, _P, _I, _I, _I, _P, _P, _I, _I, _P, _P, _I, _P]),
    "ps_bspec_scratch_bytes": (_LL, [_I]),
    "ps_bspec_triple_sums": (_I, [_I, _P, _P, _P, _I, _LL, _P, _I, _D, _P, _LL, _P, _I, _P]),
    "ps_bspec_linear_sums": (_I, [_I, _P, _P, _P, _P, _I, _LL, _P, _I, _D, _P, _LL, _P, _I, _P]),
    "ps_bspec_multi_shape_count": (_I, [_I]),
    "ps_bspec_multi_scratch_bytes": (_LL, [_I, _I, _P, _I]),
    "ps_bspec_sums_multi": (_I, [_I, _I, _P, _I, _P, _I, _P, _LL, _P]),
    "ps_bspec_multi_autotune": (_I, [_I, _I, _P, _I, _P, _LL, _I, ctypes.POINTER(_D), ctypes.POINTER(_I), _P]),
    "ps_fisher_deriv_work_bytes": (_LL, [_I, _I, _I, _I]),
    "ps_fisher_deriv_gnl_loc": (_I, [_P, _P, _I, _I, _I, _P, _P, _P, _I, _P, _I, _P, _LL, _P, _P]),
}



class BspecSegment(ctypes.Structure):
    """ps_bspec_segment: one set of stacks of the multi-segment binned-bispectrum launch."""
    _fields_ = [("mode", _I), ("nmap", _I), ("npix", _LL), ("X", _P * 4), ("triplets", _P), ("ntrip", _I),
                ("group", _I)]


class BspecGroup(ctypes.Structure):
    """ps_bspec_group: the output of the segments of one group."""
    _fields_ = [("out", _P), ("scale", _D), ("accumulate", _I)]


_lib = None


class PolySpecHIPError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PolySpecHIPError(
            "libpolyspec_hip.so not found at %s: build it with `python -m polyspec_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().ps_last_error()
        raise PolySpecHIPError("polyspec_hip error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_device(t, dtype, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise AssertionError("%s must be a device tensor" % name)
    if t.dtype != dtype:
        raise AssertionError("%s must have dtype %s (got %s)" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise AssertionError("%s must be contiguous" % name)


_workspaces = {}


def workspace(nbytes, device):
    """A cached uint8 device buffer of at least nbytes (one per device)."""
    key = torch.device(device).index or 0
    buf = _workspaces.get(key)
    if buf is None or buf.numel() < nbytes:
        _workspaces[key] = None
        buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
        _workspaces[key] = buf
    return buf


class ShtPlan:
    """HEALPix SHT plan (nside, lmax, spin, precision) on the current device.

    The plan also holds the launch shapes of its four stages per batch width (SHAPE_STAGES):
    the built-in rule unless set_shapes() or autotune() chose others.  With autotune=True the
    first transform of each batch width outside graph capture times every candidate first
    (autotune(); it synchronises the stream once)."""

    SHAPE_STAGES = ("synthesis", "ring_pow2", "ring_bluestein", "adjoint")

    def __init__(self, nside, lmax, spin=0, precision=64, autotune=False):
        lib = load()
        self.autotune_on_first_use = bool(autotune)
        self._tuned = {}
        self.nside, self.lmax, self.spin, self.precision = int(nside), int(lmax), int(spin), int(precision)
        h = ctypes.c_void_p()
        check(lib.ps_sht_plan_create(self.nside, self.lmax, self.lmax, self.spin, self.precision, ctypes.byref(h)))
        self._h = h
        npix, nlm, nc = _LL(), _LL(), _I()
        check(lib.ps_sht_plan_info(h, ctypes.byref(npix), ctypes.byref(nlm), ctypes.byref(nc)))
        self.npix, self.nlm, self.ncomp = npix.value, nlm.value, nc.value
        self.rdtype = torch.float32 if precision == 32 else torch.float64
        self.cdtype = torch.complex64 if precision == 32 else torch.complex128
        self.lmajor_multiple = lib.ps_sht_lmajor_multiple(h)
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.ps_sht_plan_destroy(h)
            self._h = None

    def work_bytes(self, nmaps, lmajor):
        return int(load().ps_sht_workspace_bytes(self._h, int(nmaps), int(lmajor)))

    # ------------------------------------------------------------ launch shapes
    def shape_candidates(self, nmaps, stage):
        """Shape ids of one stage (index or name in SHAPE_STAGES) for a batch of nmaps maps."""
        st = self.SHAPE_STAGES.index(stage) if isinstance(stage, str) else int(stage)
        lib = load()
        n = lib.ps_sht_shape_candidates(self._h, int(nmaps), st, None, 0)
        check(min(n, 0))
        ids = (_I * max(n, 1))()
        check(min(lib.ps_sht_shape_candidates(self._h, int(nmaps), st, ids, n), 0))
        return list(ids[:n])

    def shapes(self, nmaps):
        """Shapes in use for nmaps maps, one id per stage."""
        out = (_I * 4)()
        check(min(load().ps_sht_get_shapes(self._h, int(nmaps), out), 0))
        return list(out)

    def set_shapes(self, nmaps, shapes):
        """Fix the shapes for nmaps maps (a list of 4 ids, -1 = built-in rule)."""
        arr = (_I * 4)(*[int(x) for x in shapes])
        check(load().ps_sht_set_shapes(self._h, int(nmaps), arr))
        self._tuned[int(nmaps)] = None

    def autotune(self, nmaps, reps=3):
        """Time every candidate shape of every stage for nmaps maps and keep the fastest (when
        3% faster than the built-in one).  Returns {stage: (kept ms, built-in ms)} per launch."""
        lib = load()
        nb = int(lib.ps_sht_autotune_work_bytes(self._h, int(nmaps)))
        check(min(nb, 0))
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.device)
        times = (_D * 8)()
        check(lib.ps_sht_autotune(self._h, int(nmaps), ptr(ws), nb, int(reps), times, stream_handle()))
        res = {name: (times[2 * i], times[2 * i + 1]) for i, name in enumerate(self.SHAPE_STAGES)}
        self._tuned[int(nmaps)] = res
        return res

    def skip_tuning(self, nmaps):
        """Keep the built-in shapes for nmaps maps if no shapes were tuned or set for that width
        (a transform of a new width then runs without the tuning pass of autotune=True)."""
        if self.autotune_on_first_use and int(nmaps) not in self._tuned:
            self.set_shapes(nmaps, [-1] * 4)

    def _maybe_tune(self, nmaps):
        if self.autotune_on_first_use and int(nmaps) not in self._tuned and \
                not torch.cuda.is_current_stream_capturing():
            self.autotune(nmaps)

    def synthesis(self, alm, out=None, ws=None):
        """alm (nmaps, ncomp, nlm) complex -> maps (nmaps, ncomp, npix) real.  ws: optional
        uint8 device buffer of at least work_bytes(nmaps, 0) bytes (see adjoint)."""
        require_device(alm, self.cdtype, "alm")
        assert alm.dim() == 3 and alm.shape[1] == self.ncomp and alm.shape[2] == self.nlm, "Wrong alm shape"
        nm = alm.shape[0]
        self._maybe_tune(nm)
        if out is None:
            out = torch.empty((nm, self.ncomp, self.npix), dtype=self.rdtype, device=alm.device)
        require_device(out, self.rdtype, "map")
        wb = self.work_bytes(nm, 0)
        if ws is None:
            ws = workspace(wb, alm.device)
        assert ws.numel() >= wb, "workspace too small"
        check(load().ps_sht_synthesis(self._h, nm, ptr(alm), ptr(out), ptr(ws), wb, stream_handle()))
        return out

    def synthesis_grad(self, G, out=None, ws=None):
        """Spin-s synthesis of (G, C = 0): G (nmaps, nlm) complex -> maps (nmaps, 2, npix) real, the
        synthesis of (G, 0) without the curl half of the Legendre work (FP32 spin-s plans; other
        plans run the full synthesis of (G, 0))."""
        require_device(G, self.cdtype, "G")
        assert G.dim() == 2 and G.shape[1] == self.nlm, "Wrong G shape"
        assert self.ncomp == 2, "gradient-only synthesis needs a spin-s plan"
        nm = G.shape[0]
        if self.precision != 32:
            return self.synthesis(torch.stack([G, torch.zeros_like(G)], dim=1), out=out, ws=ws)
        self._maybe_tune(nm)
        if out is None:
            out = torch.empty((nm, 2, self.npix), dtype=self.rdtype, device=G.device)
        require_device(out, self.rdtype, "map")
        assert out.shape == (nm, 2, self.npix) and out.is_contiguous()
        wb = self.work_bytes(nm, 0)
        if ws is None:
            ws = workspace(wb, G.device)
        assert ws.numel() >= wb, "workspace too small"
        check(load().ps_sht_synthesis_grad(self._h, nm, ptr(G), ptr(out), ptr(ws), wb, stream_handle()))
        return out

    def synthesis_lmajor(self, alm_t, nmaps, out=None):
        """alm_t (nlm, ld) complex, columns map*ncomp+comp -> maps (nmaps, ncomp, npix)."""
        require_device(alm_t, self.cdtype, "alm_t")
        assert alm_t.dim() == 2 and alm_t.shape[0] == self.nlm
        ld = alm_t.shape[1]
        self._maybe_tune(nmaps)
        if out is None:
            out = torch.empty((nmaps, self.ncomp, self.npix), dtype=self.rdtype, device=alm_t.device)
        require_device(out, self.rdtype, "map")
        wb = self.work_bytes(nmaps, 1)
        ws = workspace(wb, alm_t.device)
        check(load().ps_sht_synthesis_lmajor(self._h, int(nmaps), ptr(alm_t), ld, ptr(out), ptr(ws), wb,
                                             stream_handle()))
        return out

    def to_lmajor(self, alm, ld=None, out=None):
        """alm (ncol, nlm) or (nmaps, ncomp, nlm) m-major -> (nlm, ld) l-major, columns >= ncol zero."""
        require_device(alm, self.cdtype, "alm")
        assert alm.shape[-1] == self.nlm, "Wrong alm shape"
        ncol = alm.numel() // self.nlm
        ld = ncol if ld is None else int(ld)
        if out is None:
            out = torch.empty((self.nlm, ld), dtype=self.cdtype, device=alm.device)
        require_device(out, self.cdtype, "alm_t")
        assert out.shape == (self.nlm, ld) and ld >= ncol, "out must be (nlm, ld >= ncol)"
        check(load().ps_sht_alm_to_lmajor(self._h, ncol, ptr(alm), ld, ptr(out), stream_handle()))
        return out

    def phase_buffer(self, nmaps):
        n = _LL()
        check(load().ps_sht_phase_bytes(self._h, int(nmaps), ctypes.byref(n)))
        return torch.empty(n.value, dtype=torch.uint8, device=self.device)

    def legendre_synthesis_lmajor(self, alm_t, nmaps, phase):
        """Legendre stage only: alm_t (nlm, ld) -> ring phases (uint8 buffer of phase_buffer())."""
        require_device(alm_t, self.cdtype, "alm_t")
        self._maybe_tune(nmaps)
        check(load().ps_sht_legendre_synthesis_lmajor(self._h, int(nmaps), ptr(alm_t), alm_t.shape[1], ptr(phase),
                                                      stream_handle()))
        return phase

    def ring_synthesis(self, phase, nmaps, out=None):
        """Ring-FFT stage only: phases -> maps (nmaps, ncomp, npix)."""
        self._maybe_tune(nmaps)
        if out is None:
            out = torch.empty((nmaps, self.ncomp, self.npix), dtype=self.rdtype, device=phase.device)
        check(load().ps_sht_ring_synthesis(self._h, int(nmaps), ptr(phase), ptr(out), stream_handle()))
        return out

    def ring_gnl_work_bytes(self, nr):
        """Bytes of the partials buffer of ring_gnl for nr radial nodes, or -1 when no staged ring
        kernel fits this plan (FP32 spin-0 plans only)."""
        if self.precision != 32 or self.spin != 0:
            return -1
        n = int(load().ps_sht_ring_gnl_work_bytes(self._h, int(nr)))
        return n if n >= 0 else -1

    def ring_gnl(self, phase, nr, w, c4, U=None, V=None, c2=0.0, out=None, accumulate=False, ws=None, inv_s2=None):
        """Ring stage fused with the gNL-loc sum: phases of 2 nr interleaved (s_r P_r, Q_r / s_r)
        columns (inv_s2 = 1 / s_r^2, None for s_r = 1) -> c4 sum_r w_r sum_p P^3 Q
        [+ c2 sum_r w_r sum_p (P^2 U + P Q V)] as an FP64 device scalar."""
        assert (U is None) == (V is None), "U and V go together"
        if inv_s2 is not None:
            require_device(inv_s2, torch.float64, "inv_s2")
            assert inv_s2.numel() == nr, "inv_s2 must have nr entries"
        for t in (U, V):
            if t is not None:
                require_device(t, torch.float64, "U/V")
                assert t.numel() == nr * self.npix, "U and V must be (nr, npix)"
        w = w if isinstance(w, torch.Tensor) else torch.as_tensor(w)
        w = w.to(self.device, torch.float64).contiguous()
        assert w.numel() == nr, "w must have nr entries"
        nb = self.ring_gnl_work_bytes(nr)
        if nb < 0:
            raise RuntimeError("ring_gnl: no staged ring kernel fits this plan")
        if ws is None:
            ws = workspace(nb, self.device)
        assert ws.numel() >= nb, "workspace too small"
        if out is None:
            out = torch.zeros((), dtype=torch.float64, device=self.device)
        check(load().ps_sht_ring_gnl(self._h, int(nr), ptr(phase), ptr(U) if U is not None else None,
                                     ptr(V) if V is not None else None, ptr(w),
                                     ptr(inv_s2) if inv_s2 is not None else None, float(c4), float(c2), ptr(ws), nb,
                                     ptr(out), int(accumulate), stream_handle()))
        return out

    def adjoint(self, maps, scale, out=None, ws=None):
        """maps (nmaps, ncomp, npix) real -> scale * adjoint (nmaps, ncomp, nlm) complex.
        ws: a uint8 device buffer of at least work_bytes(nmaps, 1) bytes to use instead of the
        shared per-device workspace (needed when transforms run concurrently on other streams)."""
        require_device(maps, self.rdtype, "map")
        assert maps.dim() == 3 and maps.shape[1] == self.ncomp and maps.shape[2] == self.npix, "Wrong map shape"
        nm = maps.shape[0]
        self._maybe_tune(nm)
        if out is None:
            out = torch.empty((nm, self.ncomp, self.nlm), dtype=self.cdtype, device=maps.device)
        require_device(out, self.cdtype, "alm")
        wb = self.work_bytes(nm, 1)
        if ws is None:
            ws = workspace(wb, maps.device)
        assert ws.numel() >= wb, "workspace too small"
        check(load().ps_sht_adjoint(self._h, nm, ptr(maps), ptr(out), float(scale), ptr(ws), wb, stream_handle()))
        return out
