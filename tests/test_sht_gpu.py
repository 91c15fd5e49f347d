"""GPU parity of the HIP SHT kernels against the FP64 oracle (same seeded inputs).

Tolerances (max |diff| / max |reference|):
  FP64 kernels: 1e-11.   FP32 kernels: 2e-6 at lmax <= 512, 5e-6 above.
"""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def rand_alm(rng, lmax, shape, lmin=0):
    ls, ms = oracle.getlm(lmax)
    a = rng.standard_normal(shape + (len(ls),)) + 1j * rng.standard_normal(shape + (len(ls),))
    a[..., ms == 0] = a[..., ms == 0].real
    a[..., ls < lmin] = 0
    return a


def relerr(x, ref):
    return float(np.abs(np.asarray(x) - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize("nside,lmax,prec,nmaps", [
    (1, 2, 64, 1), (2, 5, 64, 3), (8, 23, 64, 2), (64, 128, 64, 10), (64, 191, 64, 1),
    (8, 23, 32, 17), (64, 128, 32, 10), (128, 383, 32, 3), (256, 512, 32, 20), (64, 128, 32, 80),
    (64, 191, 32, 100), (32, 64, 32, 33)])
def test_spin0_synthesis_matches_oracle(gpu, nside, lmax, prec, nmaps):
    from polyspec_amd._abi import ShtPlan
    print(f"spin0 synthesis nside={nside} lmax={lmax} prec={prec} nmaps={nmaps}", flush=True)
    rng = np.random.default_rng(1000 + nside + lmax + nmaps)
    a = rand_alm(rng, lmax, (nmaps, 1))
    ref = oracle.synthesis(a, nside, lmax, 0)
    plan = ShtPlan(nside, lmax, 0, prec)
    out = plan.synthesis(torch.from_numpy(a).to(gpu, plan.cdtype)).cpu().numpy()
    tol = 1e-11 if prec == 64 else 2e-6
    assert relerr(out, ref) < tol


@pytest.mark.parametrize("nside,lmax,prec,nmaps", [
    (1, 2, 64, 1), (2, 5, 64, 3), (8, 23, 64, 2), (64, 128, 64, 10), (64, 191, 64, 1),
    (8, 23, 32, 5), (64, 128, 32, 10), (128, 383, 32, 3)])
def test_spin0_adjoint_matches_oracle(gpu, nside, lmax, prec, nmaps):
    from polyspec_amd._abi import ShtPlan
    print(f"spin0 adjoint nside={nside} lmax={lmax} prec={prec} nmaps={nmaps}", flush=True)
    rng = np.random.default_rng(2000 + nside + lmax + nmaps)
    m = rng.standard_normal((nmaps, 1, 12 * nside * nside))
    ref = oracle.adjoint(m, nside, lmax, 0) * 0.5
    plan = ShtPlan(nside, lmax, 0, prec)
    out = plan.adjoint(torch.from_numpy(m).to(gpu, plan.rdtype), 0.5).cpu().numpy()
    tol = 1e-11 if prec == 64 else 2e-6
    assert relerr(out, ref) < tol


@pytest.mark.parametrize("nside,lmax,prec,nmaps,spin", [
    (2, 5, 64, 1, 2), (8, 23, 64, 3, 2), (64, 191, 64, 1, 2), (8, 23, 32, 3, 2), (64, 128, 32, 2, 2),
    (128, 383, 32, 1, 2), (2, 5, 64, 1, 1), (8, 23, 64, 3, 1), (64, 191, 64, 2, 1), (64, 128, 32, 9, 1),
    (128, 255, 32, 1, 1), (8, 23, 64, 2, 3), (128, 255, 32, 21, 2), (64, 128, 32, 5, 3), (256, 512, 32, 30, 1)])
def test_spin_synthesis_and_adjoint_match_oracle(gpu, nside, lmax, prec, nmaps, spin):
    from polyspec_amd._abi import ShtPlan
    print(f"spin{spin} nside={nside} lmax={lmax} prec={prec} nmaps={nmaps}", flush=True)
    rng = np.random.default_rng(3000 + nside + lmax + nmaps + 100 * spin)
    a = rand_alm(rng, lmax, (nmaps, 2), lmin=spin)
    ref = oracle.synthesis(a, nside, lmax, spin)
    plan = ShtPlan(nside, lmax, spin, prec)
    out = plan.synthesis(torch.from_numpy(a).to(gpu, plan.cdtype)).cpu().numpy()
    tol = 1e-11 if prec == 64 else 2e-6
    assert relerr(out, ref) < tol
    m = rng.standard_normal((nmaps, 2, 12 * nside * nside))
    ref = oracle.adjoint(m, nside, lmax, spin)
    out = plan.adjoint(torch.from_numpy(m).to(gpu, plan.rdtype), 1.0).cpu().numpy()
    assert relerr(out, ref) < tol


@pytest.mark.parametrize("nside,lmax,spin", [(512, 1024, 0), (512, 1535, 2), (256, 767, 1), (1024, 2048, 0),
                                               (1024, 2048, 2)])
def test_adjointness_full_size(gpu, nside, lmax, spin):
    """<Y a, m> = <a, Y^T m>_w with w_m = 1 + [m > 0]: holds at any size (FP64 kernels)."""
    from polyspec_amd._abi import ShtPlan
    print(f"adjointness nside={nside} lmax={lmax} spin={spin}", flush=True)
    g = torch.Generator(device=gpu).manual_seed(7)
    plan = ShtPlan(nside, lmax, spin, 64)
    nc = plan.ncomp
    a = torch.randn((1, nc, plan.nlm), dtype=torch.complex128, device=gpu, generator=g)
    ls, ms = oracle.getlm(lmax)
    msk = torch.from_numpy(ms == 0).to(gpu)
    a[..., msk] = a[..., msk].real.to(torch.complex128)
    if spin:
        a[..., torch.from_numpy(ls < spin).to(gpu)] = 0
    m = torch.randn((1, nc, plan.npix), dtype=torch.float64, device=gpu, generator=g)
    Ya = plan.synthesis(a)
    YTm = plan.adjoint(m, 1.0)
    w = torch.from_numpy(1.0 + (ms > 0)).to(gpu)
    lhs = float((Ya * m).sum())
    rhs = float(((a.conj() * YTm).real * w).sum())
    assert abs(lhs - rhs) < 1e-10 * float((Ya * m).abs().sum())
