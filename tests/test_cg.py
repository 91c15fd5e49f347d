"""CG S^-1 filter: the FP64 oracle against a dense direct solve (CPU), and the
HIP CGD against the oracle on identical inputs (GPU).

Setting: a |b| < 25 deg cut, lognormal noise variance (sigma_ln = 0.5 on a
smoothed field), the test spectra of tests/estimator_setup.py."""
import numpy as np
import pytest

import oracle
from oracle.cg import OracleCG
from estimator_setup import grf_alm, spectra


def problem(nside, pol, seed=3, cut_deg=25., sigma_ln=0.5):
    rng = np.random.default_rng(seed)
    raw, tot, beam = spectra(nside, pol)
    nf = 3 if pol else 1
    th, ph = oracle.pix_angles(nside)
    lat = 90 - np.degrees(th)
    mask = (np.abs(lat) > cut_deg).astype(float) if cut_deg > 0 else np.ones_like(lat)
    L = 3 * nside - 1
    g = oracle.to_map(grf_alm(rng, {'TT': 1. / (1 + np.arange(L + 1)) ** 2}, L, False), nside, L)[0]
    g = sigma_ln * g / g.std()
    sig0 = (30 * np.pi / 10800.) ** 2 / (4 * np.pi / (12 * nside ** 2))
    ncov = np.array([sig0 * np.exp(g) * (1 if i == 0 else 2) for i in range(nf)])
    masks = np.repeat(mask[None], nf, 0)
    d = masks * (oracle.to_map(grf_alm(rng, tot, L, pol), nside, L)
                 + np.sqrt(ncov) * rng.standard_normal(ncov.shape))
    return raw, tot, beam, masks, ncov, d


def test_oracle_cg_matches_dense_solve():
    """Converged oracle CG == numpy.linalg.solve of the explicit A (Nside 4, T only; a
    well-conditioned full-sky problem so that CG reaches machine precision)."""
    nside = 4
    raw, tot, beam, mask, ncov, d = problem(nside, False, cut_deg=0., sigma_ln=0.2)
    beam = [np.asarray(beam[0])]
    o = OracleCG(nside, raw, tot, np.asarray(beam), mask, ncov, False)
    ls, ms = o.l_arr, o.m_arr
    # real parametrisation of (1, nlm) alm: Re for all, Im for m > 0
    cols = []
    basis = []
    for i in range(len(ls)):
        for part in ([0] if ms[i] == 0 else [0, 1]):
            e = np.zeros((1, len(ls)), complex)
            e[0, i] = 1. if part == 0 else 1j
            basis.append((i, part))
            Ae = o.apply_A(e)
            cols.append(np.array([(Ae[0, j].real if q == 0 else Ae[0, j].imag) for j, q in basis_all(ls, ms)]))
    A = np.array(cols).T
    b = o.get_b(d)
    bv = np.array([(b[0, j].real if q == 0 else b[0, j].imag) for j, q in basis_all(ls, ms)])
    xv = np.linalg.solve(A, bv)
    x = np.zeros((1, len(ls)), complex)
    for val, (j, q) in zip(xv, basis_all(ls, ms)):
        x[0, j] += val if q == 0 else 1j * val
    _, errs, xcg = o.solve(d, 'harmonic', nstep=300, thresh=1e-24, x0=True)
    assert errs[-1] < 1e-20
    assert np.abs(xcg - x).max() < 1e-8 * np.abs(x).max()


def basis_all(ls, ms):
    out = []
    for i in range(len(ls)):
        out.append((i, 0))
        if ms[i] > 0:
            out.append((i, 1))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("pol,prec,precond", [(False, 64, 'harmonic'), (True, 64, 'harmonic'),
                                              (True, 64, 'pseudo_inverse'), (True, 32, 'harmonic')])
def test_gpu_cg_matches_oracle(gpu, pol, prec, precond):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.weighting import Weightings
    print(f"cg pol={pol} prec={prec} precond={precond}", flush=True)
    nside = 16
    raw, tot, beam, mask, ncov, d = problem(nside, pol)
    base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=prec)
    W = Weightings(base, mask, dict(raw), noise_cov=ncov)
    o = OracleCG(nside, raw, tot, base.beam, mask, ncov, pol)
    nstep = 12
    ref, errs, _ = o.solve(d, precond, nstep=nstep, thresh=1e-30)
    from polyspec_amd.weighting import CGD
    import torch
    dd = torch.from_numpy(d).cuda().to(base.rdtype)
    a = base.to_lm(dd)
    t = W.invCtot_d
    from polyspec_amd import kernels as K
    x0 = K.lmat_apply(K.lmat_apply(a, base.lmax, mat=W.invCtot_d, post=W.beam_l_d), base.lmax, mat=W.Cl_l_d)
    cg = CGD(W, dd, x0=x0, preconditioner=precond)
    cg.run(nstep=nstep, thresh=1e-30)
    got_err = np.array([h[1] for h in cg.history])
    tol = 1e-7 if prec == 64 else 3e-3
    assert np.allclose(got_err, errs, rtol=tol, atol=0), (got_err, errs)
    out = cg.get_ivar_lm().cpu().numpy()
    assert np.abs(out - ref).max() < (1e-8 if prec == 64 else 1e-3) * np.abs(ref).max()
    # the public entry point reproduces the same solution (nstep, thresh as given)
    full = W.applySinv_optimal(d, preconditioner=precond, nstep=nstep, thresh=1e-30)
    assert np.abs(full - ref).max() < (1e-8 if prec == 64 else 1e-3) * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("pol,prec,precond", [(True, 64, 'harmonic'), (True, 64, 'pseudo_inverse'),
                                              (False, 64, 'pseudo_inverse'), (True, 32, 'harmonic')])
def test_gpu_batched_cg_matches_oracle(gpu, pol, prec, precond):
    """Four data sets solved in one batched CG: every system against the oracle's own solve."""
    from polyspec_amd.base import PolySpec
    from polyspec_amd.weighting import Weightings
    print(f"batched cg pol={pol} prec={prec} precond={precond}", flush=True)
    nside = 16
    raw, tot, beam, mask, ncov, d = problem(nside, pol)
    ds = np.stack([d] + [problem(nside, pol, seed=10 + s)[5] for s in range(3)])
    base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=prec)
    W = Weightings(base, mask, dict(raw), noise_cov=ncov)
    o = OracleCG(nside, raw, tot, base.beam, mask, ncov, pol)
    nstep = 12
    out = W.applySinv_optimal_batch(ds, preconditioner=precond, nstep=nstep, thresh=1e-30)
    assert out.shape == (4, 3 if pol else 1, (base.lmax + 1) * (base.lmax + 2) // 2)
    tol = 1e-8 if prec == 64 else 1e-3
    for s in range(4):
        ref, errs, _ = o.solve(ds[s], precond, nstep=nstep, thresh=1e-30)
        assert np.abs(out[s] - ref).max() < tol * np.abs(ref).max(), s
        got = np.array([h[1][s] for h in W.last_cg.history])
        assert np.allclose(got, errs, rtol=1e-7 if prec == 64 else 3e-3, atol=0), (s, got, errs)


@pytest.mark.gpu
def test_gpu_batched_cg_stops_each_system_like_a_single_solve(gpu):
    """With a threshold, each system of the batch stops at the iteration a separate solve stops
    at and returns the same solution (converged systems stay frozen while the others run)."""
    from polyspec_amd.base import PolySpec
    from polyspec_amd.weighting import Weightings
    nside = 16
    raw, tot, beam, mask, ncov, d = problem(nside, True)
    ds = np.stack([d] + [problem(nside, True, seed=20 + s)[5] for s in range(4)])
    base = PolySpec(nside, dict(tot), beam=beam, pol=True, precision=64)
    W = Weightings(base, mask, dict(raw), noise_cov=ncov)
    out = W.applySinv_optimal_batch(ds, preconditioner='harmonic', nstep=200, thresh=1e-5)
    iters = W.last_cg.iterations.copy()
    assert (W.last_cg.err < 1e-5).all() and iters.max() < 200
    print("iterations per system", iters, flush=True)
    for s in range(len(ds)):
        single = W.applySinv_optimal(ds[s], preconditioner='harmonic', nstep=200, thresh=1e-5)
        assert W.last_cg.i == iters[s], (s, W.last_cg.i, iters)
        # the two runs differ only in the rounding of their inner products (blocking of the
        # FP64 reductions), which CG amplifies over its iterations
        assert np.abs(out[s] - single).max() < 1e-7 * np.abs(single).max(), s


@pytest.mark.gpu
def test_gpu_batched_cg_update_freezes_inactive_systems(gpu):
    import torch
    from polyspec_amd import kernels as K
    g = torch.Generator(device=gpu).manual_seed(5)
    shape = (3, 3, 1000)
    p, Ap, x, r = (torch.randn(shape, dtype=torch.complex128, device=gpu, generator=g) for _ in range(4))
    x0, r0, p0 = x.clone(), r.clone(), p.clone()
    rz = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device=gpu)
    pAp = torch.tensor([4.0, 5.0, 6.0], dtype=torch.float64, device=gpu)
    active = torch.tensor([1, 0, 1], dtype=torch.int32, device=gpu)
    K.cg_update_xr_batch(rz, pAp, active, p, Ap, x, r)
    z = torch.randn(shape, dtype=torch.complex128, device=gpu, generator=g)
    K.cg_update_p_batch(pAp, rz, active, z, p)
    for b in (0, 2):
        al = float(rz[b] / pAp[b])
        assert torch.allclose(x[b], x0[b] + al * p0[b], rtol=1e-14, atol=1e-14)
        assert torch.allclose(r[b], r0[b] - al * Ap[b], rtol=1e-14, atol=1e-14)
        assert torch.allclose(p[b], z[b] + float(pAp[b] / rz[b]) * p0[b], rtol=1e-14, atol=1e-14)
    assert torch.equal(x[1], x0[1]) and torch.equal(r[1], r0[1]) and torch.equal(p[1], p0[1])
