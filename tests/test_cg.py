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
