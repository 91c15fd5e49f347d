"""Learned preconditioner for the CG S^-1 filter, and the flexible CG that uses it.

The reference's CG (polyspec/weighting.py:217-367) preconditions with the
isotropic harmonic inverse Sigma = (C^-1 + B^2/N_iso)^-1 (weighting.py:75-78,
304-305), exact only for a full sky with white noise.  Over the observed sky at
large scales (B^2/N >> C^-1) the true inverse is ~ N(x)/B^2 with the local noise
variance N(x), so Sigma errs by the factor N(x)/N_iso, which a lognormal noise
map spreads over an order of magnitude.  A small network learns the correction
on a coarse grid:

  M(r) = Sigma r + P^dag [ g * f_theta(features(r)) ] s,

with P the synthesis of the l <= lmax_lo part of a harmonic vector on a coarse
HEALPix grid (Nside_lo, default 128), P^dag its adjoint times the coarse pixel
area (a quadrature inverse of P for l <= 2 Nside_lo), g the coarse mask clipped
to [0, 1] (inside the mask the data carry no information and a local correction
only harms the conditioning), and f_theta a 4-layer, 64-channel
9-point-stencil convolutional network over the coarse pixels (each pixel and its 8 HEALPix
neighbours, healpix.stencil9).  Its input
channels are P r, P Sigma r and P C r, each divided by its rms, plus two static
channels, the coarse mask and the normalised log noise variance; its output is
in units of s = rms(P Sigma r), so M(a r) = a M(r) for a > 0.

Training pairs are exact and free: every plain harmonic CG iteration applies A to
its search direction p_k, and (A p_k, p_k) is a residual/solution pair of the
same operator, spanning the Krylov space the solver works in.  The target of a
pair (r, e) is g * P(e - Sigma r) / s.  The network runs on the device through
libpolyspec_hip.so (im2col gather + hand-written GEMM kernels, backward passes,
Adam); FP32 parameters and activations, FP64 loss.

M is nonlinear, so FlexibleCGD uses the flexible (Polak-Ribiere) beta
(Notay 2000, FCG(1)).  Plain CG stays the judge -- convergence is always
measured with the reference statistic r^T Sigma r / b0^T Sigma b0, whatever
preconditioner is in use -- and the fallback: a non-finite or non-positive
r^T M r, or a judge statistic that stops improving, drops the network and the
solver continues as plain harmonic CG from the current iterate.
"""
import time

import numpy as np
import torch

from . import kernels as K
from ._abi import ShtPlan, check, load, ptr, stream_handle
from .healpix import stencil9
from .weighting import CGD


# --------------------------------------------------------------- network
class StencilNet:
    """Device parameters, forward/backward and Adam of the stencil network.

    Layer i: Y = act(im2col9(X) W_i + b_i), W_i (9 cin, cout) row-major, ReLU on
    every layer but the last."""

    def __init__(self, nbr, rev, channels=(5, 64, 64, 64, 3), seed=0, device=None, gate=None, bf16_inference=True):
        self.device = device or torch.device('cuda', torch.cuda.current_device())
        # forward passes that keep nothing for backward (inference) run on bf16 matrix cores with
        # the stencil gather fused (ps_nn_conv9_bf16); training keeps the FP32 kernels
        self.bf16_inference = bf16_inference
        self._packed, self._packed_key, self._set_count = None, None, 0
        # optional per-pixel factor (npix,) applied to the last layer's output
        self.gate = None if gate is None else torch.as_tensor(gate).to(self.device, torch.float32).contiguous()
        self.nbr = torch.from_numpy(np.ascontiguousarray(nbr)).to(self.device)
        self.rev = torch.from_numpy(np.ascontiguousarray(rev)).to(self.device)
        self.rw = int(rev.shape[1])
        self.npix = int(nbr.shape[0])
        self.channels = [int(c) for c in channels]
        rng = np.random.default_rng(seed)
        self.params = []
        for i in range(len(channels) - 1):
            cin, cout = self.channels[i], self.channels[i + 1]
            last = i == len(channels) - 2
            gain = (0.1 if last else 1.) * np.sqrt(2. / (9 * cin))
            W = rng.standard_normal((9 * cin, cout)) * gain
            self.params.append([torch.from_numpy(W).float().to(self.device).contiguous(),
                                torch.zeros(cout, dtype=torch.float32, device=self.device)])
        self.m = [[torch.zeros_like(W), torch.zeros_like(b)] for W, b in self.params]
        self.v = [[torch.zeros_like(W), torch.zeros_like(b)] for W, b in self.params]
        self.t = 0
        self._work = {}

    def set_params(self, params):
        for (W, b), (Wn, bn) in zip(self.params, params):
            W.copy_(torch.as_tensor(np.asarray(Wn), dtype=torch.float32))
            b.copy_(torch.as_tensor(np.asarray(bn), dtype=torch.float32))
        self._set_count += 1

    @staticmethod
    def _cin_bf16(cin):
        """Channels of a layer's bf16 input rows: cin rounded up to 16, 32 or 64."""
        return 16 if cin <= 16 else 32 if cin <= 32 else 64

    def _bf16_weights(self):
        """Every layer's W in bf16 as ps_nn_conv9_bf16 reads it, (32 ceil(cout/32), 9 cp) with cp the
        padded input channels (zero rows there); rebuilt whenever a parameter changed (Adam steps,
        set_params, or an in-place torch operation, seen through the tensors' version counters)."""
        key = (self.t, self._set_count) + tuple(t._version for pr in self.params for t in pr)
        if self._packed_key != key:
            packed = []
            for i, (W, b) in enumerate(self.params):
                cin, cout = self.channels[i], self.channels[i + 1]
                cp = self._cin_bf16(cin)
                rows = 32 if cout <= 32 else 64
                Wp = torch.zeros((rows, 9, cp), dtype=torch.float32, device=self.device)
                Wp[:cout, :, :cin] = W.view(9, cin, cout).permute(2, 0, 1)
                packed.append(Wp.reshape(rows, 9 * cp).to(torch.bfloat16).contiguous())
            self._packed, self._packed_key = packed, key
        return self._packed

    def _bf16_ok(self):
        """The bf16 kernels take hidden widths of 16, 32 or 64, inputs of at most 64 channels and
        outputs of at most 64."""
        ch = self.channels
        return ch[0] <= 64 and all(c in (16, 32, 64) for c in ch[1:-1]) and ch[-1] <= 64

    def _forward_bf16(self, X):
        """Inference on bf16 matrix cores: the input padded to 16/32/64 bf16 channels, hidden
        activations kept in bf16, the last layer's output in FP32."""
        L = load()
        B = X.shape[0]
        c0 = self.channels[0]
        cp = self._cin_bf16(c0)
        h = torch.zeros((B, self.npix, cp), dtype=torch.bfloat16, device=self.device)
        h[..., :c0] = X
        nl = len(self.params)
        st = stream_handle()
        for i, ((W, b), Wp) in enumerate(zip(self.params, self._bf16_weights())):
            cin, cout = self.channels[i], self.channels[i + 1]
            last = i == nl - 1
            cnext = cout if last else self._cin_bf16(cout)
            assert cnext == cout or last, "hidden widths must be 16, 32 or 64 for the bf16 path"
            y = torch.empty((B, self.npix, cout), dtype=torch.float32 if last else torch.bfloat16, device=self.device)
            gate = self.gate if (last and self.gate is not None) else None
            check(L.ps_nn_conv9_bf16(ptr(h), ptr(self.nbr), B, self.npix, self._cin_bf16(cin), ptr(Wp), ptr(b), cout,
                                     int(not last), ptr(gate) if gate is not None else None, ptr(y), int(not last), st))
            h = y
        return h

    def _buf(self, key, n):
        t = self._work.get(key)
        if t is None or t.numel() < n:
            t = torch.empty(int(n), dtype=torch.float32, device=self.device)
            self._work[key] = t
        return t

    def forward(self, X, keep=False, bf16=None):
        """X (B, npix, C0) float32 -> (B, npix, C_last).  keep=True stores what backward needs (FP32
        kernels); otherwise the bf16 matrix-core form runs unless bf16=False (default:
        bf16_inference)."""
        L = load()
        assert X.dtype == torch.float32 and X.dim() == 3 and X.shape[1] == self.npix \
            and X.shape[2] == self.channels[0], "StencilNet: bad input shape"
        if not keep and (self.bf16_inference if bf16 is None else bf16) and self._bf16_ok():
            return self._forward_bf16(X)
        B = X.shape[0]
        M = B * self.npix
        h = X.contiguous()
        self._cols, self._outs = [], []
        nl = len(self.params)
        st = stream_handle()
        for i, (W, b) in enumerate(self.params):
            cin, cout = self.channels[i], self.channels[i + 1]
            if keep:
                col = torch.empty((M, 9 * cin), dtype=torch.float32, device=self.device)
            else:
                col = self._buf('col', M * 9 * cin)[:M * 9 * cin].view(M, 9 * cin)
            check(L.ps_nn_im2col9(ptr(h), ptr(self.nbr), B, self.npix, cin, ptr(col), st))
            y = torch.empty((B, self.npix, cout), dtype=torch.float32, device=self.device)
            check(L.ps_nn_gemm(0, 0, M, cout, 9 * cin, ptr(col), 9 * cin, ptr(W), cout, ptr(y), cout, ptr(b),
                               int(i < nl - 1), 1, None, st))
            if keep:
                self._cols.append(col)
                self._outs.append(y)
            h = y
        if self.gate is not None:
            check(L.ps_nn_pixel_scale(B, self.npix, self.channels[-1], ptr(self.gate), ptr(h), ptr(h), st))
        return h

    def backward(self, dY):
        """Parameter gradients [(dW, db)] of the last forward(keep=True) for output gradient dY."""
        L = load()
        B = dY.shape[0]
        M = B * self.npix
        nl = len(self.params)
        grads = [None] * nl
        d = dY.contiguous()
        st = stream_handle()
        if self.gate is not None:
            d = d.clone()
            check(L.ps_nn_pixel_scale(B, self.npix, self.channels[-1], ptr(self.gate), ptr(d), ptr(d), st))
        for i in range(nl - 1, -1, -1):
            W, b = self.params[i]
            cin, cout = self.channels[i], self.channels[i + 1]
            if i < nl - 1:
                check(L.ps_nn_relu_backward(d.numel(), ptr(self._outs[i]), ptr(d), st))
            dW = torch.empty_like(W)
            ks = max(1, min(256, M // 2048))
            work = self._buf('split', ks * 9 * cin * cout)
            check(L.ps_nn_gemm(1, 0, 9 * cin, cout, M, ptr(self._cols[i]), 9 * cin, ptr(d), cout, ptr(dW), cout,
                               None, 0, ks, ptr(work), st))
            db = torch.empty_like(b)
            cw = self._buf('colsum', int(L.ps_nn_colsum_work_floats(M, cout)))
            check(L.ps_nn_colsum(M, cout, ptr(d), ptr(cw), ptr(db), st))
            grads[i] = (dW, db)
            if i > 0:
                G = self._buf('dcol', M * 9 * cin)[:M * 9 * cin].view(M, 9 * cin)
                check(L.ps_nn_gemm(0, 1, M, 9 * cin, cout, ptr(d), cout, ptr(W), cout, ptr(G), 9 * cin, None, 0, 1,
                                   None, st))
                dX = torch.empty((B, self.npix, cin), dtype=torch.float32, device=self.device)
                check(L.ps_nn_col2im9(ptr(G), ptr(self.rev), self.rw, B, self.npix, cin, ptr(dX), st))
                d = dX
        self._cols, self._outs = [], []
        return grads

    def mse(self, Y, T, scale, with_grad=True):
        """loss = sum_b scale_b |Y_b - T_b|^2 (FP64 device scalar) and dloss/dY."""
        L = load()
        B = Y.shape[0]
        s = torch.as_tensor(scale, dtype=torch.float64).to(self.device).reshape(B).contiguous()
        dY = torch.empty_like(Y) if with_grad else None
        work = torch.empty(512, dtype=torch.float64, device=self.device)
        loss = torch.empty((), dtype=torch.float64, device=self.device)
        check(L.ps_nn_mse(Y[0].numel(), B, ptr(Y), ptr(T.contiguous()), ptr(s), ptr(dY) if with_grad else None,
                          ptr(work), ptr(loss), stream_handle()))
        return loss, dY

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = load()
        self.t += 1
        st = stream_handle()
        for (W, b), (gW, gb), (mW, mb), (vW, vb) in zip(self.params, grads, self.m, self.v):
            check(L.ps_nn_adam(W.numel(), ptr(W), ptr(gW), ptr(mW), ptr(vW), lr, b1, b2, eps, self.t, st))
            check(L.ps_nn_adam(b.numel(), ptr(b), ptr(gb), ptr(mb), ptr(vb), lr, b1, b2, eps, self.t, st))


# ------------------------------------------------------ learned preconditioner
def lm_select(L_in, L_out):
    """Indices of the l <= L_out entries of an m-major (healpy) alm array of lmax L_in."""
    return np.concatenate([np.arange(m * (2 * L_in + 1 - m) // 2 + m, m * (2 * L_in + 1 - m) // 2 + L_out + 1)
                           for m in range(L_out + 1)])


class LearnedPreconditioner:
    """M(r) = Sigma r + P^dag f(features(r)) s on a coarse grid (module docstring)."""

    def __init__(self, weightings, nside_lo=128, lmax_lo=None, channels=64, nlayers=4, seed=0):
        W = weightings
        self.W = W
        self.base = W.base
        self.nf = W.nf
        self.L = self.base.lmax
        self.nside_lo = int(nside_lo)
        self.Llo = int(min(self.L, 2 * nside_lo if lmax_lo is None else lmax_lo))
        dev = self.base.device
        with torch.cuda.device(dev):
            tune = getattr(self.base, 'autotune', False)
            self.p0 = ShtPlan(nside_lo, self.Llo, 0, self.base.precision, autotune=tune)
            self.p2 = ShtPlan(nside_lo, self.Llo, 2, self.base.precision, autotune=tune) if self.nf == 3 else None
        self.npix_lo = 12 * self.nside_lo ** 2
        self.sel = torch.from_numpy(lm_select(self.L, self.Llo)).to(dev)
        nbr, rev = stencil9(self.nside_lo)
        self.nin = 4 * self.nf + 2
        ch = [self.nin] + [channels] * (nlayers - 1) + [self.nf]
        # static channels: coarse mask and normalised log noise variance (harmonic downgrades)
        nc = W.noise_cov[0]
        pos = nc > 0
        lv = np.log(np.where(pos, nc, np.median(nc[pos]) if pos.any() else 1.))
        lv = (lv - lv[pos].mean()) / (lv[pos].std() + 1e-30) if pos.any() else 0. * lv
        hi = np.stack([W.mask_arr[0], lv])
        self.static = self._down_maps(hi).to(torch.float32).contiguous()  # (npix_lo, 2)
        self.gate = self.static[:, 0].clamp(0., 1.).contiguous()
        # local noise contrast N(x)/N_iso - 1 per field on the coarse grid
        iso = W.Nl_iso / W.pixel_area
        self.dN = (self._down_maps(W.noise_cov) / torch.from_numpy(iso).to(dev)[None, :] - 1.).to(self.base.rdtype)
        self.net = StencilNet(nbr, rev, ch, seed=seed, device=dev, gate=self.gate)
        self.enabled = True
        self.n_calls = 0
        self.time = 0.

    def _down_maps(self, maps):
        """(k, Npix) host maps on the CG grid -> (npix_lo, k) coarse maps via l <= Llo."""
        pm = self.base.plan(self.Llo, 0)
        m = torch.from_numpy(np.ascontiguousarray(maps)).to(self.base.device, self.base.rdtype)
        a = pm.adjoint(m.reshape(-1, 1, m.shape[-1]), self.base.A_pix)
        return self.p0.synthesis(a)[:, 0, :].t()

    def project(self, v):
        """(k, nf, nlm(L)) harmonic vectors -> (npix_lo, k*nf) coarse maps of their l <= Llo part."""
        k = v.shape[0]
        a = v.index_select(-1, self.sel)
        out = self.p0.synthesis(a[:, 0:1].contiguous())
        if self.nf == 3:
            out = torch.cat([out, self.p2.synthesis(a[:, 1:3].contiguous())], dim=1)
        return out.reshape(k * self.nf, self.npix_lo).t()

    def back_project(self, maps):
        """(npix_lo, nf) coarse maps -> (nf, nlm(Llo)) harmonic: adjoint times the coarse pixel area."""
        m = maps.t().to(self.base.rdtype).contiguous()
        sc = 4 * np.pi / self.npix_lo
        a0 = self.p0.adjoint(m[0:1].view(1, 1, -1), sc)[0]
        if self.nf == 1:
            return a0
        a2 = self.p2.adjoint(m[1:3].reshape(1, 2, -1), sc)[0]
        return torch.cat([a0, a2], dim=0)

    def _dyn(self, r, Sr):
        """Dynamic features (npix_lo, 4 nf) float32 -- P r, P Sigma r, P C r and the local
        noise-contrast product (N/N_iso - 1) P(B Sigma r / N_iso), each over its rms -- and
        the output scale s = rms(P Sigma r)."""
        Cr = K.lmat_apply(r, self.L, mat=self.W.Cl_l_d)
        bSr = K.lmat_apply(Sr, self.L, post=self.W.bn_d)
        maps = self.project(torch.stack([r, Sr, Cr, bSr])).reshape(self.npix_lo, 4, self.nf).clone()
        maps[:, 3] *= self.dN
        ms = torch.sqrt(torch.mean(maps ** 2, dim=(0, 2))).clamp_min(1e-300)
        feats = (maps / ms[None, :, None]).reshape(self.npix_lo, 4 * self.nf)
        return feats.to(torch.float32), ms[1]

    def __call__(self, r):
        return self.apply(r)[0]

    def apply(self, r):
        """(M r, Sigma r): the preconditioned residual and the harmonic part alone (the judge's
        Sigma r, so that flexible CG does not apply Sigma twice)."""
        t0 = time.time()
        Sr = K.lmat_apply(r, self.L, mat=self.W.Sigma_l_d)
        if not self.enabled:
            return Sr, Sr
        feats, s = self._dyn(r, Sr)
        X = torch.cat([feats, self.static], dim=-1)[None].contiguous()
        Y = self.net.forward(X)[0]
        corr = self.back_project(Y) * s
        z = Sr.clone()
        z.index_add_(-1, self.sel, corr.to(Sr.dtype))
        self.n_calls += 1
        self.time += time.time() - t0
        return z, Sr

    # ----------------------------------------------------------- training
    def collect(self, data_maps, max_pairs=2000, iters=60, verb=False):
        """Exact residual/solution pairs from plain harmonic CG solves: at iteration k the
        solver applies A to its search direction p_k, and (r, e) = (A p_k, p_k).  Per pair
        keep the dynamic features of r and the target g * P(e - Sigma r) / s.  Returns
        device tensors X (n, npix_lo, 4 nf) and T (n, npix_lo, nf), float32."""
        Xs, Ts = [], []
        g = self.gate[:, None]
        for d in data_maps:
            d = torch.as_tensor(d).to(self.base.device, self.base.rdtype).reshape(self.nf, -1).contiguous()
            cg = CGD(self.W, d, x0=self.W.wiener_start(d), preconditioner='harmonic')
            cg.keep_Ap = True
            n0 = len(Xs)
            for _ in range(iters):
                cg.step()
                r, e = cg.last_Ap, cg.last_p
                Sr = K.lmat_apply(r, self.L, mat=self.W.Sigma_l_d)
                f, s = self._dyn(r, Sr)
                t = self.project((e - Sr)[None]).to(torch.float32) * g / s.to(torch.float32)
                Xs.append(f)
                Ts.append(t)
                if len(Xs) >= max_pairs:
                    break
            if verb:
                print("collected %d pairs (%d from a solve reaching %.1e)" % (len(Xs), len(Xs) - n0, cg.err),
                      flush=True)
            if len(Xs) >= max_pairs:
                break
        return torch.stack(Xs), torch.stack(Ts)

    def train(self, X, T, epochs=4, batch=4, lr=1e-3, seed=0, time_limit=None, verb=False):
        """Adam on the mean relative squared error |f(X_b) - T_b|^2 / |T_b|^2 with a cosine
        learning-rate decay.  Returns the per-epoch mean relative loss."""
        n = X.shape[0]
        norms = (T.double() ** 2).sum(dim=(1, 2)).clamp_min(1e-300)
        g = torch.Generator().manual_seed(seed)
        nb = max(1, n // batch)
        total = epochs * nb
        hist, it, t0 = [], 0, time.time()
        static = self.static[None].expand(batch, -1, -1)
        stop = False
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(X.device)
            tot = torch.zeros((), dtype=torch.float64, device=X.device)
            for bi in range(nb):
                idx = perm[bi * batch:(bi + 1) * batch]
                xb = torch.cat([X.index_select(0, idx), static[:len(idx)]], dim=-1).contiguous()
                tb = T.index_select(0, idx).contiguous()
                yb = self.net.forward(xb, keep=True)
                loss, dY = self.net.mse(yb, tb, 1. / (norms.index_select(0, idx) * len(idx)))
                grads = self.net.backward(dY)
                lr_i = lr * 0.5 * (1 + np.cos(np.pi * it / total))
                self.net.adam_step(grads, lr=lr_i)
                tot += loss * len(idx)
                it += 1
                if time_limit is not None and time.time() - t0 > time_limit:
                    stop = True
                    break
            hist.append(float(tot) / ((bi + 1) * batch))
            if verb:
                print("epoch %d: mean relative loss %.4f (%.1f s)" % (ep, hist[-1], time.time() - t0), flush=True)
            if stop:
                break
        return hist

    def relative_loss(self, X, T, batch=8):
        """Mean |f(X) - T|^2 / |T|^2 over a held-out set."""
        out = []
        for i in range(0, X.shape[0], batch):
            xb = torch.cat([X[i:i + batch], self.static[None].expand(min(batch, X.shape[0] - i), -1, -1)], dim=-1)
            y = self.net.forward(xb.contiguous())
            t = T[i:i + batch]
            out.append(((y.double() - t.double()) ** 2).sum(dim=(1, 2)) / (t.double() ** 2).sum(dim=(1, 2)))
        return float(torch.cat(out).mean())


# ------------------------------------------------------------ flexible CG
class FlexibleCGD(CGD):
    """CG with a (possibly nonlinear) preconditioner: Polak-Ribiere beta, convergence judged
    by the reference statistic with the harmonic Sigma, fallback to plain harmonic CG."""

    def __init__(self, weightings, d, precond, x0=[], stall=10, verb=False):
        self.precond = precond
        self.stall = stall
        self.fallback_at = None
        super().__init__(weightings, d, x0=x0, preconditioner='harmonic', verb=verb)
        self.best = self.err
        self.nbad = 0
        z = self.precond(self.r)
        rz = self.dot(self.r, z)
        if self._usable(rz):
            self.z, self.rz, self.p = z, rz, z.clone()
        else:
            self._fall_back()

    @staticmethod
    def _usable(rz):
        v = float(rz)
        return np.isfinite(v) and v > 0

    def _fall_back(self):
        """Continue as plain harmonic CG from the current iterate."""
        if getattr(self.precond, 'enabled', False):
            self.precond.enabled = False
        self.fallback_at = self.i
        self.z = K.lmat_apply(self.r, self.L, mat=self.weightings.Sigma_l_d)
        self.rz = self.dot(self.r, self.z)
        self.p = self.z.clone()
        self.nbad = 0

    def step(self):
        if self.fallback_at is not None:
            return CGD.step(self)
        Ap = self._apply_A(self.p)
        pAp = self.dot(self.p, Ap)
        r_old = self.r.clone()
        K.cg_update_xr(self.rz, pAp, self.p, Ap, self.x, self.r)
        del Ap
        self.i += 1
        # M r and the judge's Sigma r, then every scalar the step decides on in one transfer
        if hasattr(self.precond, 'apply'):
            z, judge = self.precond.apply(self.r)
        else:
            z, judge = self.precond(self.r), K.lmat_apply(self.r, self.L, mat=self.weightings.Sigma_l_d)
        rz_new = self.dot(self.r, z)
        num = rz_new - self.dot(r_old, z)  # Polak-Ribiere (flexible) numerator
        err, rr, rz_h = torch.stack([self.dot(self.r, judge) / self.rz0, self.dot(self.r, self.r),
                                     rz_new]).cpu().tolist()
        self.err = err
        self.history.append((self.i, self.err, float(np.sqrt(rr)) / self.bnorm))
        if not np.isfinite(self.err):
            raise FloatingPointError("flexible CG: non-finite residual")
        if self.err < self.best:
            self.best, self.nbad = self.err, 0
        else:
            self.nbad += 1
        if not self._usable(rz_h) or self.nbad >= self.stall:
            self._fall_back()
            return
        K.cg_update_p(num, self.rz, z, self.p)
        self.rz = rz_new
        self.z = z
