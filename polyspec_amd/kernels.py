"""Device-tensor wrappers of the radial-node kernels in libpolyspec_hip.so.

Each function mirrors one Cython routine of the reference
(polyspec/cython/tauNL_utils.pyx, ideal_fisher.pyx) but takes torch device
tensors and enqueues on torch's current stream.  Scalar results are returned
as 0-d float64 device tensors (no host synchronisation inside).
"""
import torch

from . import _abi
from ._abi import check, load, ptr, stream_handle

_scratch = {}


def _prec(t):
    if t.dtype in (torch.float32, torch.complex64):
        return 32
    if t.dtype in (torch.float64, torch.complex128):
        return 64
    raise AssertionError("unsupported dtype %s" % t.dtype)


def _dev_f64(x, device):
    if isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64 and x.is_contiguous():
        return x
    return torch.as_tensor(x, dtype=torch.float64).to(device).contiguous()


def reduce_scratch(device):
    key = torch.device(device).index or 0
    s = _scratch.get(key)
    if s is None:
        n = int(load().ps_reduce_scratch_bytes()) // 8
        s = torch.empty(n, dtype=torch.float64, device=device)
        _scratch[key] = s
    return s


def _chk(*ts):
    for t in ts:
        if t is not None:
            assert isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous(), "device, contiguous tensors only"


def _nlm(lmax):
    return (int(lmax) + 1) * (int(lmax) + 2) // 2


def _same(ref, *ts, what="operands"):
    """Every tensor of ts has the shape and dtype of ref (the kernels read them with ref's extent)."""
    for t in ts:
        if t is not None:
            assert t.shape == ref.shape and t.dtype == ref.dtype, \
                "%s must share shape and dtype: %s %s vs %s %s" % (what, tuple(ref.shape), ref.dtype, tuple(t.shape),
                                                                  t.dtype)


def _stack(t, nr, npix, what):
    """t holds an (nr, npix) stack (any leading shape of nr * npix values)."""
    assert t.numel() == nr * npix, "%s must be an (%d, %d) stack, got %s" % (what, nr, npix, tuple(t.shape))


def _table(f, lmax, nr, what):
    """Radial table (>= lmax+1, nslot, nr)."""
    assert f.dim() == 3 and f.shape[0] > lmax and f.shape[2] == nr, \
        "%s must be (>= %d, nslot, %d), got %s" % (what, lmax + 1, nr, tuple(f.shape))


def weight_alm_lmajor(f, h, lmax, lmin, weight, out_t, col_offset=0, npol=None):
    """apply_fl_weights (tauNL_utils.pyx:80-98) into l-major synthesis input columns.

    f: (lmax+1, nslot, nr) real, h: (npol, nlm(lmax)) complex, out_t: (nlm, ld) complex.
    """
    _chk(f, h, out_t)
    nr = f.shape[2]
    npol = min(f.shape[1], h.shape[0], 2) if npol is None else npol
    _table(f, lmax, nr, "f")
    assert h.dim() == 2 and h.shape[0] >= npol and h.shape[1] == _nlm(lmax), "h must be (npol, nlm(lmax))"
    assert out_t.dim() == 2 and out_t.shape[0] >= _nlm(lmax) and out_t.shape[1] >= col_offset + nr, \
        "out_t must be (>= nlm(lmax), >= col_offset + nr)"
    check(load().ps_weight_alm_lmajor(_prec(h), ptr(f), f.shape[1], nr, ptr(h), npol, int(lmax), int(lmin),
                                      float(weight), out_t.shape[1], int(col_offset), ptr(out_t), stream_handle()))
    return out_t


def weight_alm_pq_lmajor(fP, fQ, h, lmax, lmin, wP, wQ, out_t, col_offset=0, npol=None, interleaved=False):
    """The P and Q weightings of one field (apply_fl_weights with fP, wP and with fQ, wQ) into the
    l-major columns col_offset + [0, nr) and col_offset + [nr, 2 nr) (interleaved: P_r at
    col_offset + 2r, Q_r at col_offset + 2r + 1), one launch.  fP and fQ: (>= lmax+1, nslot, nr)
    with the same nslot and nr (their l ranges may differ)."""
    _chk(fP, fQ, h, out_t)
    assert fP.shape[1:] == fQ.shape[1:] and fP.dtype == fQ.dtype, "P and Q tables must share their layout"
    assert min(fP.shape[0], fQ.shape[0]) > lmax, "the tables must cover l <= lmax"
    nr = fP.shape[2]
    npol = min(fP.shape[1], h.shape[0], 2) if npol is None else npol
    assert h.dim() == 2 and h.shape[0] >= npol and h.shape[1] == _nlm(lmax), "h must be (npol, nlm(lmax))"
    assert out_t.dim() == 2 and out_t.shape[0] >= _nlm(lmax) and out_t.shape[1] >= col_offset + 2 * nr, \
        "out_t must be (>= nlm(lmax), >= col_offset + 2 nr)"
    fn = load().ps_weight_alm_pq_interleaved_lmajor if interleaved else load().ps_weight_alm_pq_lmajor
    check(fn(_prec(h), ptr(fP), ptr(fQ), fP.shape[1], nr, ptr(h), npol, int(lmax), int(lmin), float(wP), float(wQ),
             out_t.shape[1], int(col_offset), ptr(out_t), stream_handle()))
    return out_t


def gnl_loc_fused_sum(P, Q, w, c4=1.0, U=None, V=None, c2=0.0, out=None, accumulate=False):
    """c4 * gnl_loc_sum(w, P, Q) [+ c2 * sum_r w_r sum_p (P^2 U + P Q V)] as a device scalar (U, V FP64)."""
    _chk(P, Q, U, V)
    assert (U is None) == (V is None), "U and V go together"
    assert U is None or (U.dtype == torch.float64 and V.dtype == torch.float64), "U, V must be float64"
    assert P.dim() >= 2, "P must be an (nr, npix) stack"
    nr, npix = P.shape[-2], P.shape[-1]
    _same(P, Q, what="P and Q")
    if U is not None:
        _stack(U, nr, npix, "U")
        _stack(V, nr, npix, "V")
    w = _dev_f64(w, P.device)
    assert w.numel() == nr, "w must have nr = %d entries" % nr
    if out is None:
        out = torch.zeros((), dtype=torch.float64, device=P.device)
    check(load().ps_gnl_loc_fused_sum(_prec(P), ptr(P), ptr(Q), ptr(U) if U is not None else None,
                                      ptr(V) if V is not None else None, ptr(w), nr, npix, float(c4), float(c2),
                                      ptr(reduce_scratch(P.device)), ptr(out), int(accumulate), stream_handle()))
    return out


def gnl_loc_sum(w, P, Q):
    """tauNL_utils.gnl_loc_sum (tauNL_utils.pyx:260-270)."""
    return gnl_loc_fused_sum(P, Q, w, 1.0)


def gnl_loc_disc_sum(w, P1, P2, Q1, Q2, out=None, accumulate=False):
    """tauNL_utils.gnl_loc_disc_sum (tauNL_utils.pyx:275-285)."""
    _chk(P1, P2, Q1, Q2)
    assert P1.dim() >= 2, "P1 must be an (nr, npix) stack"
    _same(P1, P2, Q1, Q2, what="P1, P2, Q1 and Q2")
    w = _dev_f64(w, P1.device)
    assert w.numel() == P1.shape[-2], "w must have nr = %d entries" % P1.shape[-2]
    if out is None:
        out = torch.zeros((), dtype=torch.float64, device=P1.device)
    check(load().ps_gnl_loc_disc_sum(_prec(P1), ptr(P1), ptr(P2), ptr(Q1), ptr(Q2), ptr(w), P1.shape[-2],
                                     P1.shape[-1], ptr(reduce_scratch(P1.device)), ptr(out), int(accumulate),
                                     stream_handle()))
    return out


def mc_accumulate(P, Q, U, V):
    """U += sum_s P_s Q_s, V += sum_s P_s^2 over the leading (simulation) axis of P, Q (U, V FP64)."""
    _chk(P, Q, U, V)
    assert U.dtype == torch.float64 and V.dtype == torch.float64, "Monte-Carlo sums U, V must be float64"
    _same(P, Q, what="P and Q")
    assert U.shape == V.shape, "U and V must share their shape"
    n = U.numel()
    nsim = P.numel() // n
    assert n > 0 and nsim * n == P.numel() and tuple(P.shape[P.dim() - U.dim():]) == tuple(U.shape), \
        "P must stack whole copies of U's shape: %s vs %s" % (tuple(P.shape), tuple(U.shape))
    check(load().ps_mc_accumulate(_prec(P), ptr(P), ptr(Q), nsim, n, ptr(U), ptr(V), stream_handle()))


def _products(mode, a, b, c=None, d=None, scale=1.0, out=None):
    _chk(a, b, c, d)
    assert mode in (0, 2) or d is not None, "this product needs four operands"
    assert mode in (0,) or c is not None, "this product needs a third operand"
    for t in (b, c, d):  # the kernel reads a.numel() values of every operand
        assert t is None or (t.numel() == a.numel() and t.dtype == a.dtype), \
            "map operands must share size and dtype: %s %s vs %s %s" % (tuple(a.shape), a.dtype, tuple(t.shape), t.dtype)
    if out is None:
        out = torch.empty_like(a)
    _chk(out)
    assert out.numel() == a.numel() and out.dtype == a.dtype, "out must match the operands"
    check(load().ps_pixel_products(_prec(a), mode, ptr(a), ptr(b), ptr(c) if c is not None else None,
                                   ptr(d) if d is not None else None, a.numel(), float(scale), ptr(out),
                                   stream_handle()))
    return out


def multiplyRR(a, b, out=None):
    """ideal_fisher.multiplyRR (ideal_fisher.pyx:100-108)."""
    return _products(0, a, b, out=out)


def multiplyRR_sym(a, b, c, d, out=None):
    """ideal_fisher.multiplyRR_sym (ideal_fisher.pyx:112-120): a b + c d."""
    return _products(1, a, b, c, d, out=out)


def multiply(a, b, c, scale=1.0, out=None):
    """tauNL_utils.multiply (tauNL_utils.pyx:103-111): a b c."""
    return _products(2, a, b, c, scale=scale, out=out)


def multiply_asym(a1A, a1B, a3A, a3B, scale=1.0, out=None):
    """tauNL_utils.multiply_asym (tauNL_utils.pyx:116-124): A (2 B C + A D)."""
    return _products(3, a1A, a1B, a3A, a3B, scale=scale, out=out)


def radial_sum(lm_map, w, f, lmax, lmin, npol, out=None, scale=1.0, accumulate=False):
    """tauNL_utils.radial_sum (tauNL_utils.pyx:220-235): (nr, nlm) -> (npol, nlm)."""
    _chk(lm_map, f)
    assert lm_map.dim() == 2 and lm_map.shape[1] == _nlm(lmax), "lm_map must be (nr, nlm(lmax))"
    nr = lm_map.shape[0]
    _table(f, lmax, nr, "f")
    assert npol <= f.shape[1], "npol exceeds the table's slots"
    w = _dev_f64(w, lm_map.device)
    assert w.numel() == nr, "w must have nr = %d entries" % nr
    if out is None:
        out = torch.zeros((npol, lm_map.shape[-1]), dtype=lm_map.dtype, device=lm_map.device)
    _chk(out)
    assert out.dtype == lm_map.dtype and out.numel() == npol * lm_map.shape[1], "out must be (npol, nlm(lmax))"
    check(load().ps_radial_sum(_prec(lm_map), ptr(lm_map), lm_map.shape[0], int(lmax), int(lmin), ptr(f),
                               f.shape[1], int(npol), ptr(w), float(scale), int(accumulate), ptr(out),
                               stream_handle()))
    return out


def radial_scatter(lm_map, w, f, lmax, lmin, npol, out=None, scale=1.0, accumulate=False):
    """Per-node terms of radial_sum, ideal_fisher.integrate_pq_deriv (ideal_fisher.pyx:395-407):
    (nr, nlm) -> (nr, npol, nlm)."""
    _chk(lm_map, f)
    assert lm_map.dim() == 2 and lm_map.shape[1] == _nlm(lmax), "lm_map must be (nr, nlm(lmax))"
    nr = lm_map.shape[0]
    _table(f, lmax, nr, "f")
    assert npol <= f.shape[1], "npol exceeds the table's slots"
    w = _dev_f64(w, lm_map.device)
    assert w.numel() == nr, "w must have nr = %d entries" % nr
    if out is None:
        out = torch.zeros((nr, npol, lm_map.shape[1]), dtype=lm_map.dtype, device=lm_map.device)
    _chk(out)
    assert out.dtype == lm_map.dtype and out.numel() == nr * npol * lm_map.shape[1], "out must be (nr, npol, nlm(lmax))"
    check(load().ps_radial_scatter(_prec(lm_map), ptr(lm_map), nr, int(lmax), int(lmin), ptr(f), f.shape[1], int(npol),
                                   ptr(w), float(scale), int(accumulate), ptr(out), stream_handle()))
    return out


def tau_sum_n0(A, B, FLL, w, Lmaxs, Lmin, Lmax, n1=0, out=None, accumulate=False):
    """tauNL_utils.tau_sum_n0 (tauNL_utils.pyx:604-646); A, B: (nmu, nr, nlm(Lmaxs))."""
    _chk(A, B)
    _same(A, B, what="A and B")
    assert A.dim() == 3 and A.shape[2] == _nlm(Lmaxs), "A, B must be (nmu, nr, nlm(Lmaxs))"
    nr = A.shape[1]
    FLL = _dev_f64(FLL, A.device)
    w = _dev_f64(w, A.device)
    assert FLL.numel() == (int(Lmaxs) + 1) * nr * nr, "FLL must be (Lmaxs+1, nr, nr)"
    assert w.numel() == nr, "w must have nr = %d entries" % nr
    assert 0 <= Lmin and Lmax <= Lmaxs, "L range outside the A maps"
    if out is None:
        out = torch.zeros((), dtype=torch.float64, device=A.device)
    check(load().ps_tau_sum_n0(_prec(A), ptr(A), ptr(B), A.shape[0], A.shape[1], int(Lmaxs), int(Lmin), int(Lmax),
                               ptr(FLL), ptr(w), int(n1), ptr(reduce_scratch(A.device)), ptr(out), int(accumulate),
                               stream_handle()))
    return out


def tau_sum_lmajor(At, Bt, FLL, w, nr, nmu, Lmaxs, Lmin, Lmax, n1=0, out=None, accumulate=False, M=None,
                   shiftA=False, shiftB=False):
    """Sum over mu < nmu of tau_sum_n0 (tauNL_utils.pyx:604-646) of the pairs (A_mu, B_mu), with
    A, B l-major (nlm(Lmaxs), ld), column mu * nr + j; Bt may be At (tau(A, A)).  With M
    (nlm(Lmaxs), nr) l-major: the pairs (A_mu - M, B_mu - M), shifted on the sides selected by
    shiftA / shiftB."""
    _chk(At, Bt, M)
    if shiftA or shiftB:
        assert M is not None and M.dtype == At.dtype and M.numel() == _nlm(Lmaxs) * nr, "M must be (nlm(Lmaxs), nr)"
    assert At.dim() == 2 and At.shape[0] == _nlm(Lmaxs), "At must be (nlm(Lmaxs), ld)"
    _same(At, Bt, what="At and Bt")
    assert At.shape[1] >= nmu * nr and 1 <= nr <= 64, "ld must cover nmu * nr columns (nr <= 64)"
    FLL = _dev_f64(FLL, At.device)
    w = _dev_f64(w, At.device)
    assert FLL.numel() == (int(Lmaxs) + 1) * nr * nr, "FLL must be (Lmaxs+1, nr, nr)"
    assert w.numel() == nr, "w must have nr = %d entries" % nr
    assert 0 <= Lmin and Lmax <= Lmaxs, "L range outside the A maps"
    if out is None:
        out = torch.zeros((), dtype=torch.float64, device=At.device)
    check(load().ps_tau_sum_lmajor_shift(_prec(At), ptr(At), ptr(Bt), At.shape[1], int(nmu), int(nr), int(Lmaxs),
                                         int(Lmin), int(Lmax), ptr(FLL), ptr(w), int(n1),
                                         ptr(M) if M is not None else None, int(bool(shiftA)), int(bool(shiftB)),
                                         ptr(reduce_scratch(At.device)), ptr(out), int(accumulate), stream_handle()))
    return out


def outer_product(Qa, Qb, out=None, accumulate=False):
    """ideal_fisher.outer_product (ideal_fisher.pyx:423-459), sym=False."""
    _chk(Qa, Qb)
    assert Qa.dim() == 3 and Qb.dim() == 3 and Qa.shape[0] == 4 and Qb.shape[0] == 4, "Q must be (4, nt, nin)"
    assert Qa.shape[2] == Qb.shape[2] and Qa.dtype == Qb.dtype, "Qa and Qb must share nin and dtype"
    if out is None:
        out = torch.zeros((Qa.shape[1], Qb.shape[1]), dtype=torch.float64, device=Qa.device)
    _chk(out)
    assert out.dtype == torch.float64 and out.shape == (Qa.shape[1], Qb.shape[1]), "fish must be (nta, ntb) float64"
    check(load().ps_outer_product(_prec(Qa), ptr(Qa), ptr(Qb), Qa.shape[1], Qb.shape[1], Qa.shape[2],
                                  ptr(reduce_scratch(Qa.device)), ptr(out), int(accumulate), stream_handle()))
    return out


def convolve_F_lmajor(PQ, FLL, w, Lmaxs, Lmin, Lmax, out_t):
    """convolve_F_n000 (tauNL_utils.pyx:966-992) into l-major synthesis columns 0..nr-1 of out_t."""
    _chk(PQ, out_t)
    assert PQ.dim() == 2 and PQ.shape[1] == _nlm(Lmaxs), "PQ must be (nr, nlm(Lmaxs))"
    nr = PQ.shape[0]
    FLL = _dev_f64(FLL, PQ.device)
    w = _dev_f64(w, PQ.device)
    assert FLL.numel() == (int(Lmaxs) + 1) * nr * nr, "FLL must be (Lmaxs+1, nr, nr)"
    assert w.numel() == nr, "w must have nr = %d entries" % nr
    assert out_t.dim() == 2 and out_t.shape[0] >= _nlm(Lmaxs) and out_t.shape[1] >= nr and \
        out_t.dtype == PQ.dtype, "out_t must be (>= nlm(Lmaxs), >= nr) of PQ's dtype"
    check(load().ps_convolve_F_lmajor(_prec(PQ), ptr(PQ), PQ.shape[0], int(Lmaxs), int(Lmin), int(Lmax), ptr(FLL),
                                      ptr(w), out_t.shape[1], ptr(out_t), stream_handle()))
    return out_t


def lmat_apply(x, lmax, mat=None, pre=None, post=None, out=None, alpha=1.0, accumulate=False):
    """out[a] (+)= alpha post[a,l] sum_b mat[a,b,l] pre[b,l] x[b] on (nf, nlm) vectors."""
    _chk(x)
    nf = x.shape[0]
    dev = x.device
    m = _dev_f64(mat, dev) if mat is not None else None
    pr = _dev_f64(pre, dev) if pre is not None else None
    po = _dev_f64(post, dev) if post is not None else None
    if out is None:
        out = torch.empty_like(x)
    _chk(out)
    check(load().ps_lmat_apply(_prec(x), nf, int(lmax), ptr(m) if m is not None else None,
                               ptr(pr) if pr is not None else None, ptr(po) if po is not None else None, ptr(x),
                               x.shape[-1], ptr(out), out.shape[-1], float(alpha), int(accumulate), stream_handle()))
    return out


_dot_scratch = {}


def mdot(x, y, lmax, out=None):
    """CGD.dot (weighting.py:255) as a 0-d float64 device tensor."""
    _chk(x, y)
    key = x.device.index or 0
    s = _dot_scratch.get(key)
    if s is None:
        s = torch.empty(int(load().ps_mdot_scratch_bytes()) // 8, dtype=torch.float64, device=x.device)
        _dot_scratch[key] = s
    if out is None:
        out = torch.empty((), dtype=torch.float64, device=x.device)
    nf = x.numel() // (((lmax + 1) * (lmax + 2)) // 2)
    check(load().ps_mdot(_prec(x), nf, int(lmax), ptr(x), ptr(y), ptr(s), ptr(out), stream_handle()))
    return out


def cg_update_xr(rz, pAp, p, Ap, x, r):
    _chk(rz, pAp, p, Ap, x, r)
    check(load().ps_cg_update_xr(_prec(x), x.numel(), ptr(rz), ptr(pAp), ptr(p), ptr(Ap), ptr(x), ptr(r),
                                 stream_handle()))


def cg_update_p(num, den, z, p):
    _chk(num, den, z, p)
    check(load().ps_cg_update_p(_prec(p), p.numel(), ptr(num), ptr(den), ptr(z), ptr(p), stream_handle()))


# ------------------------------------------------------------ batched CG (many maps in lock step)
def multiplyRR_batch(a, b, out=None):
    """multiplyRR of every stack of a (B, ...) with the shared b (a[0].shape)."""
    _chk(a, b)
    nb = a.shape[0]
    assert b.numel() * nb == a.numel(), "b must have the shape of one stack of a"
    if out is None:
        out = torch.empty_like(a)
    check(load().ps_pixel_products_batch(_prec(a), 0, nb, ptr(a), ptr(b), None, None, b.numel(), 1.0, ptr(out),
                                         stream_handle()))
    return out


def multiplyRR_sym_batch(a, b, c, d, out=None):
    """a[s] b + c[s] d for every stack s of a and c (B, ...) with the shared b, d (a[0].shape)."""
    _chk(a, b, c, d)
    nb = a.shape[0]
    assert c.shape == a.shape and c.dtype == a.dtype, "a and c must share shape and dtype"
    for t in (b, d):
        assert t.numel() * nb == a.numel() and t.dtype == a.dtype, "b, d must have the shape of one stack of a"
    if out is None:
        out = torch.empty_like(a)
    _chk(out)
    assert out.shape == a.shape and out.dtype == a.dtype, "out must match a"
    check(load().ps_pixel_products_batch(_prec(a), 4, nb, ptr(a), ptr(b), ptr(c), ptr(d), b.numel(), 1.0, ptr(out),
                                         stream_handle()))
    return out


def lmat_apply_batch(x, lmax, mat=None, pre=None, post=None, out=None, alpha=1.0, accumulate=False):
    """lmat_apply on every vector of x (B, nf, nlm)."""
    _chk(x)
    assert x.dim() == 3 and x.shape[2] == ((lmax + 1) * (lmax + 2)) // 2
    dev = x.device
    m = _dev_f64(mat, dev) if mat is not None else None
    pr = _dev_f64(pre, dev) if pre is not None else None
    po = _dev_f64(post, dev) if post is not None else None
    if out is None:
        out = torch.empty_like(x)
    _chk(out)
    check(load().ps_lmat_apply_batch(_prec(x), x.shape[0], x.shape[1], int(lmax), ptr(m) if m is not None else None,
                                     ptr(pr) if pr is not None else None, ptr(po) if po is not None else None,
                                     ptr(x), ptr(out), float(alpha), int(accumulate), stream_handle()))
    return out


_dot_scratch_b = {}


def mdot_batch(x, y, lmax, out=None):
    """CGD.dot of each system of a batch, x, y (B, nf, nlm) -> (B,) float64 device tensor."""
    _chk(x, y)
    nb, nf = x.shape[0], x.shape[1]
    key = (x.device.index or 0, nb)
    s = _dot_scratch_b.get(key)
    if s is None:
        s = torch.empty(int(load().ps_mdot_batch_scratch_bytes(nb)) // 8, dtype=torch.float64, device=x.device)
        _dot_scratch_b[key] = s
    if out is None:
        out = torch.empty(nb, dtype=torch.float64, device=x.device)
    check(load().ps_mdot_batch(_prec(x), nb, nf, int(lmax), ptr(x), ptr(y), ptr(s), ptr(out), stream_handle()))
    return out


def cg_update_xr_batch(rz, pAp, active, p, Ap, x, r):
    _chk(rz, pAp, p, Ap, x, r)
    nb = p.shape[0]
    check(load().ps_cg_update_xr_batch(_prec(x), nb, x[0].numel(), ptr(rz), ptr(pAp), ptr(active), ptr(p), ptr(Ap),
                                       ptr(x), ptr(r), stream_handle()))


def cg_update_p_batch(num, den, active, z, p):
    _chk(num, den, z, p)
    nb = p.shape[0]
    check(load().ps_cg_update_p_batch(_prec(p), nb, p[0].numel(), ptr(num), ptr(den), ptr(active), ptr(z), ptr(p),
                                      stream_handle()))


__all__ = ["lmat_apply", "mdot", "cg_update_xr", "cg_update_p", "lmat_apply_batch", "mdot_batch",
           "cg_update_xr_batch", "cg_update_p_batch", "multiplyRR_batch", "multiplyRR_sym_batch", "convolve_F_lmajor", "weight_alm_lmajor", "weight_alm_pq_lmajor", "gnl_loc_fused_sum", "gnl_loc_sum", "gnl_loc_disc_sum", "mc_accumulate",
           "multiplyRR", "multiplyRR_sym", "multiply", "multiply_asym", "radial_sum", "tau_sum_n0", "tau_sum_lmajor",
           "outer_product", "_abi"]
