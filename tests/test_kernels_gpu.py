"""GPU parity of the radial-node and CG vector kernels against the FP64 oracle
loop nests / numpy definitions on seeded inputs.  FP64 kernels: 1e-12 relative;
FP32: 1e-5 relative (FP64 accumulation, FP32 storage)."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def T(x, dt):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda().to(dt)


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("lmax", [0, 1, 7, 30, 200])
def test_lmat_apply(gpu, prec, lmax):
    from polyspec_amd import kernels as K
    rng = np.random.default_rng(lmax)
    ls, ms = oracle.getlm(lmax)
    nf = 3
    x = rng.standard_normal((nf, len(ls))) + 1j * rng.standard_normal((nf, len(ls)))
    M = rng.standard_normal((nf, nf, lmax + 1))
    pre = rng.random((nf, lmax + 1))
    post = rng.random((nf, lmax + 1))
    cdt = torch.complex128 if prec == 64 else torch.complex64
    y = K.lmat_apply(T(x, cdt), lmax, mat=M, pre=pre, post=post, alpha=0.5)
    ref = 0.5 * post[:, ls] * np.einsum('abl,bl->al', M[:, :, ls], pre[:, ls] * x)
    assert rel(y, ref) < (1e-13 if prec == 64 else 1e-5)
    d = K.mdot(T(x, cdt), T(x * 0.3 + 1j, cdt), lmax)
    dref = np.sum((x * np.conj(x * 0.3 + 1j)).real * (1. + (ms > 0)))
    assert abs(float(d) - dref) < (1e-12 if prec == 64 else 1e-5) * np.sum(np.abs(x) * np.abs(x * 0.3 + 1j) * 2)


@pytest.mark.parametrize("prec", [64, 32])
def test_weight_radial_tau_outer(gpu, prec):
    from polyspec_amd import kernels as K
    rng = np.random.default_rng(5)
    lmax, lmin, nr = 40, 2, 6
    ls, ms = oracle.getlm(lmax)
    rdt = torch.float64 if prec == 64 else torch.float32
    cdt = torch.complex128 if prec == 64 else torch.complex64
    tol = 1e-12 if prec == 64 else 2e-5
    f = rng.standard_normal((lmax + 2, 3, nr))
    h = rng.standard_normal((2, len(ls))) + 1j * rng.standard_normal((2, len(ls)))
    ld = 16
    out = torch.zeros((len(ls), ld), dtype=cdt, device='cuda')
    K.weight_alm_lmajor(T(f, rdt), T(h, cdt), lmax, lmin, 0.7, out, col_offset=3)
    sel = ls >= lmin
    ref = np.zeros((nr, len(ls)), complex)
    ref[:, sel] = oracle.apply_fl_weights(f, h[:, sel], ls[sel], 0.7)
    got = out.cpu().numpy()
    assert rel(got[:, 3:3 + nr].T, ref) < tol
    assert np.all(got[:, :3] == 0) and np.all(got[:, 3 + nr:] == 0)
    # both weightings of one field in one launch equal the two single launches bit for bit
    f2 = rng.standard_normal((lmax + 2, 3, nr))
    two = torch.full((len(ls), 2 * nr + 4), float("nan"), dtype=cdt, device='cuda')
    K.weight_alm_lmajor(T(f, rdt), T(h, cdt), lmax, lmin, 0.7, two, col_offset=2)
    K.weight_alm_lmajor(T(f2, rdt), T(h, cdt), lmax, lmin, -1.3, two, col_offset=2 + nr)
    one = torch.full_like(two, float("nan"))
    K.weight_alm_pq_lmajor(T(f, rdt), T(f2, rdt), T(h, cdt), lmax, lmin, 0.7, -1.3, one, col_offset=2)
    assert torch.equal(one[:, 2:2 + 2 * nr], two[:, 2:2 + 2 * nr])
    assert torch.isnan(one[:, :2].real).all() and torch.isnan(one[:, 2 + 2 * nr:].real).all()
    # the four-column form (nr, ld and col_offset multiples of 4) agrees with the one-column form
    n4 = 8
    g1, g2 = rng.standard_normal((2, lmax + 2, 3, n4))
    two = torch.full((len(ls), 2 * n4 + 8), float("nan"), dtype=cdt, device='cuda')
    K.weight_alm_lmajor(T(g1, rdt), T(h, cdt), lmax, lmin, 0.7, two, col_offset=4)
    K.weight_alm_lmajor(T(g2, rdt), T(h, cdt), lmax, lmin, -1.3, two, col_offset=4 + n4)
    one = torch.full_like(two, float("nan"))
    K.weight_alm_pq_lmajor(T(g1, rdt), T(g2, rdt), T(h, cdt), lmax, lmin, 0.7, -1.3, one, col_offset=4)
    assert torch.equal(one[:, 4:4 + 2 * n4], two[:, 4:4 + 2 * n4])
    assert torch.isnan(one[:, :4].real).all() and torch.isnan(one[:, 4 + 2 * n4:].real).all()

    lm = rng.standard_normal((nr, len(ls))) + 1j * rng.standard_normal((nr, len(ls)))
    w = rng.random(nr)
    rs = K.radial_sum(T(lm, cdt), w, T(f, rdt), lmax, lmin, 3, scale=2.0)
    ref = np.zeros((3, len(ls)), complex)
    ref[:, sel] = 2.0 * oracle.radial_sum(lm[:, sel], w, f, ls[sel])
    assert rel(rs, ref) < tol

    # tau_sum_n0 over L in [Lmin, Lmax] with the (nmu, nr, nlm) layout
    Lmaxs, Lmin, Lmax = 30, 1, 25
    Ls, Ms = oracle.getlm(Lmaxs)
    A = rng.standard_normal((1, nr, len(Ls))) + 1j * rng.standard_normal((1, nr, len(Ls)))
    B = rng.standard_normal((1, nr, len(Ls))) + 1j * rng.standard_normal((1, nr, len(Ls)))
    F = rng.standard_normal((Lmaxs + 1, nr, nr))
    w2 = w.copy()
    w2[1] = 0.
    got = float(K.tau_sum_n0(T(A, cdt), T(B, cdt), F, w2, Lmaxs, Lmin, Lmax))
    ref = oracle.tau_sum_n0(np.transpose(A, (0, 2, 1)), np.transpose(B, (0, 2, 1)), F, w2, Ls, Ms, Lmin, Lmax)
    assert abs(got - ref) < (1e-11 if prec == 64 else 1e-5) * np.abs(A).sum() * np.abs(B).max() * np.abs(F).max()

    # convolve_F into l-major columns
    buf = torch.zeros((len(Ls), 8), dtype=cdt, device='cuda')
    K.convolve_F_lmajor(T(A[0], cdt), F, w2, Lmaxs, Lmin, Lmax, buf)
    Li = np.where((Ls >= Lmin) & (Ls <= Lmax), np.arange(len(Ls)), -1).astype(np.int32)
    ref = oracle.convolve_F_n000(np.ascontiguousarray(A[0].T), F, w2, Ls, Li, len(Ls))
    assert rel(buf.cpu().numpy()[:, :nr], ref) < tol

    Qa = rng.standard_normal((4, 2, 300)) + 1j * rng.standard_normal((4, 2, 300))
    Qb = rng.standard_normal((4, 2, 300)) + 1j * rng.standard_normal((4, 2, 300))
    assert rel(K.outer_product(T(Qa, cdt), T(Qb, cdt)), oracle.outer_product(Qa, Qb)) < tol


@pytest.mark.parametrize("prec", [64, 32])
def test_pixel_kernels(gpu, prec):
    from polyspec_amd import kernels as K
    rng = np.random.default_rng(9)
    nr, npix = 5, 3000
    rdt = torch.float64 if prec == 64 else torch.float32
    tol = 1e-12 if prec == 64 else 1e-5
    a, b, c, d = (rng.standard_normal((nr, npix)) for _ in range(4))
    w = rng.random(nr)
    ta, tb, tc, td = (T(x, rdt) for x in (a, b, c, d))
    assert rel(K.multiplyRR(ta, tb), a * b) < tol
    assert rel(K.multiplyRR_sym(ta, tb, tc, td), a * b + c * d) < tol
    assert rel(K.multiply(ta, tb, tc), a * b * c) < tol
    assert rel(K.multiply_asym(ta, tb, tc, td), a * (2 * b * c + a * d)) < tol
    s = float(K.gnl_loc_sum(w, ta, tb))
    assert abs(s - oracle.gnl_loc_sum(w, a, b)) < tol * np.sum(w[:, None] * np.abs(a) ** 3 * np.abs(b))
    s = float(K.gnl_loc_disc_sum(w, ta, tb, tc, td))
    ref = oracle.gnl_loc_disc_sum(w, a, b, c, d)
    assert abs(s - ref) < tol * np.sum(w[:, None] * (a * a * np.abs(b * d) + b * b * np.abs(a * c)))
    U = torch.zeros_like(ta, dtype=torch.float64)  # Monte-Carlo sums are FP64 in both precisions
    V = torch.zeros_like(ta, dtype=torch.float64)
    K.mc_accumulate(torch.stack([tb, tc]).contiguous(), torch.stack([td, ta]).contiguous(), U, V)
    assert rel(U, b * d + c * a) < tol and rel(V, b * b + c * c) < tol
    f = float(K.gnl_loc_fused_sum(ta, tb, w, 0.3, U, V, -0.7))
    ref = 0.3 * oracle.gnl_loc_sum(w, a, b) - 0.7 * (oracle.gnl_loc_disc_sum(w, a, b, b, d) +
                                                      oracle.gnl_loc_disc_sum(w, a, c, b, a))
    assert abs(f - ref) < 10 * tol * np.sum(w[:, None] * np.abs(a) ** 2 * (np.abs(a * b) + np.abs(U.cpu().numpy())
                                                                           + np.abs(V.cpu().numpy())))
    # a pixel count that is not a multiple of 4 takes the one-pixel-per-step kernel
    a3, b3 = a[:, :2999].copy(), b[:, :2999].copy()
    s = float(K.gnl_loc_sum(w, T(a3, rdt), T(b3, rdt)))
    assert abs(s - oracle.gnl_loc_sum(w, a3, b3)) < tol * np.sum(w[:, None] * np.abs(a3) ** 3 * np.abs(b3))


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("nr,nmu,same", [(3, 1, True), (16, 3, False), (17, 2, True), (40, 2, False), (64, 1, True)])
def test_tau_sum_lmajor_matches_oracle(gpu, prec, nr, nmu, same):
    """ps_tau_sum_lmajor (per-L Gram products on FP64 matrix cores, l-major A/B batched over mu)
    against the sum over mu of the oracle's tau_sum_n0 (tauNL_utils.pyx:604-646); the columns
    beyond nmu * nr hold NaN, which must never be read."""
    from polyspec_amd import kernels as K
    from polyspec_amd._abi import ShtPlan
    print(f"tau lmajor prec={prec} nr={nr} nmu={nmu} same={same}", flush=True)
    rng = np.random.default_rng(nr * 10 + nmu)
    Lmaxs, Lmin, Lmax = 23, 2, 21
    Ls, Ms = oracle.getlm(Lmaxs)
    nlm = len(Ls)
    cdt = torch.complex128 if prec == 64 else torch.complex64
    A = rng.standard_normal((nmu, nr, nlm)) + 1j * rng.standard_normal((nmu, nr, nlm))
    B = A.copy() if same else rng.standard_normal((nmu, nr, nlm)) + 1j * rng.standard_normal((nmu, nr, nlm))
    F = rng.standard_normal((Lmaxs + 1, nr, nr))
    F = F + F.transpose(0, 2, 1)
    w = rng.random(nr)
    w[nr // 2] = 0.
    # oracle: its (nmu, nLM, nr) layout over Lmin..Lmax rows, i.e. L >= Lmin of 0..Lmax
    keep = Ls <= Lmax
    sel = Ls[keep] >= Lmin
    ref = 0.
    for mu in range(nmu):
        a = A[mu][:, keep][:, sel].T[None].copy()
        b = B[mu][:, keep][:, sel].T[None].copy()
        ref += oracle.tau_sum_n0(a, b, F, w, Ls[keep][sel].astype(np.int32), Ms[keep][sel].astype(np.int32), Lmin,
                                 Lmax)
    ld = nmu * nr + 5
    plan = ShtPlan(8, Lmaxs, 0, prec)
    def lmaj(X):
        t = torch.full((nlm, ld), float('nan'), dtype=cdt, device=gpu)
        t[:, :nmu * nr] = T(X.reshape(nmu * nr, nlm).T, cdt)
        return t
    At = lmaj(A)
    Bt = At if same else lmaj(B)
    got = float(K.tau_sum_lmajor(At, Bt, F, w, nr, nmu, Lmaxs, Lmin, Lmax))
    assert abs(got - ref) < (1e-12 if prec == 64 else 2e-6) * np.sum(np.abs(F).max() * np.abs(A).max() ** 2 * nr ** 2), \
        (got, ref)
    assert abs(got - ref) < (1e-11 if prec == 64 else 1e-5) * abs(ref) + 1e-300, (got, ref)
    # the mean-field shift: pairs (A_mu - M, B_mu - M) on the selected sides, against the oracle on
    # explicitly shifted maps (the shift is applied in FP64 inside the kernel)
    Mh = rng.standard_normal((nr, nlm)) + 1j * rng.standard_normal((nr, nlm))
    Mh = T(Mh, cdt).cpu().numpy()  # the values the kernel sees
    Ah, Bh = At[:, :nmu * nr].T.cpu().numpy().reshape(nmu, nr, nlm), Bt[:, :nmu * nr].T.cpu().numpy().reshape(nmu, nr, nlm)
    Mt = T(Mh.T, cdt).contiguous()
    for sa, sb in ([(True, True), (False, True)] if same else [(True, False), (True, True)]):
        ref_s = 0.
        for mu in range(nmu):
            a = (Ah[mu].astype(np.complex128) - (Mh if sa else 0))[:, keep][:, sel].T[None].copy()
            b = (Bh[mu].astype(np.complex128) - (Mh if sb else 0))[:, keep][:, sel].T[None].copy()
            ref_s += oracle.tau_sum_n0(a, b, F, w, Ls[keep][sel].astype(np.int32), Ms[keep][sel].astype(np.int32),
                                       Lmin, Lmax)
        got_s = float(K.tau_sum_lmajor(At, Bt, F, w, nr, nmu, Lmaxs, Lmin, Lmax, M=Mt, shiftA=sa, shiftB=sb))
        assert abs(got_s - ref_s) < 1e-12 * np.sum(np.abs(F).max() * 4 * np.abs(A).max() ** 2 * nr ** 2), \
            (sa, sb, got_s, ref_s)
    # the transpose into that layout, and the batched symmetric product of simulation maps
    Am = T(A.reshape(nmu * nr, nlm), cdt)
    lt = plan.to_lmajor(Am, ld=ld)
    assert torch.equal(lt[:, :nmu * nr], At[:, :nmu * nr]) and torch.all(lt[:, nmu * nr:] == 0)
    rdt = torch.float64 if prec == 64 else torch.float32
    Qs, Ps = T(rng.standard_normal((3, nr, 50)), rdt), T(rng.standard_normal((3, nr, 50)), rdt)
    P, Q = T(rng.standard_normal((nr, 50)), rdt), T(rng.standard_normal((nr, 50)), rdt)
    prod = K.multiplyRR_sym_batch(Qs, P, Ps, Q)
    want = Qs.cpu().numpy() * P.cpu().numpy() + Ps.cpu().numpy() * Q.cpu().numpy()
    assert rel(prod, want) < (1e-15 if prec == 64 else 1e-6)
