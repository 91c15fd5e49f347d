"""Launch shapes of the SHT stages: every candidate shape of every stage, and the shapes the
autotuner keeps, against the FP64 oracle on the same seeded inputs.

Tolerances (max |diff| / max |reference|): FP64 kernels 1e-11, FP32 kernels 2e-6 (as for the
built-in shapes).
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


def _case(nside, lmax, spin, nmaps, seed):
    rng = np.random.default_rng(seed)
    nc = 1 if spin == 0 else 2
    a = rand_alm(rng, lmax, (nmaps, nc), lmin=spin)
    m = rng.standard_normal((nmaps, nc, 12 * nside * nside))
    return a, m, oracle.synthesis(a, nside, lmax, spin), oracle.adjoint(m, nside, lmax, spin)


@pytest.mark.parametrize("nside,lmax,prec,spin,nmaps", [
    (32, 64, 32, 0, 1), (32, 64, 32, 0, 20), (32, 64, 64, 0, 3), (32, 64, 32, 2, 5), (32, 64, 64, 2, 1),
    (16, 47, 32, 1, 3), (64, 128, 32, 0, 2), (64, 128, 64, 2, 2), (256, 512, 32, 0, 17), (256, 512, 32, 2, 9),
    (64, 128, 32, 0, 80), (32, 95, 32, 0, 100), (128, 200, 32, 0, 40),
    # FP32 spin-s batches of two tiles and more (the wide spin-s synthesis at 2, 3 and 4 tiles)
    (32, 64, 32, 1, 9), (64, 128, 32, 2, 40),
    # the spin-s matrix-core adjoint with two chunks of 8 maps per block, the last one partial
    (64, 128, 32, 2, 21), (256, 512, 32, 2, 17),
    # FP64 batches (the FP64 matrix-core adjoint; ring slices and partial sums from Nside 256)
    (32, 64, 64, 0, 20), (64, 128, 64, 2, 9), (256, 512, 64, 0, 12), (256, 512, 64, 2, 5)])
def test_every_candidate_shape_matches_oracle(gpu, nside, lmax, prec, spin, nmaps):
    from polyspec_amd._abi import ShtPlan
    a, m, ref_map, ref_alm = _case(nside, lmax, spin, nmaps, 4000 + nside + lmax + nmaps + 100 * spin)
    plan = ShtPlan(nside, lmax, spin, prec)
    ta = torch.from_numpy(a).to(gpu, plan.cdtype)
    tm = torch.from_numpy(m).to(gpu, plan.rdtype)
    tol = 1e-11 if prec == 64 else 2e-6
    for stage in range(4):
        cands = plan.shape_candidates(nmaps, stage)
        assert len(cands) >= 1
        assert plan.shapes(nmaps)[stage] in cands  # the built-in choice is a candidate
        for c in cands:
            print(f"nside={nside} lmax={lmax} prec={prec} spin={spin} nmaps={nmaps} stage={stage} shape={c}",
                  flush=True)
            shapes = [-1] * 4
            shapes[stage] = c
            plan.set_shapes(nmaps, shapes)
            assert plan.shapes(nmaps)[stage] == c
            assert relerr(plan.synthesis(ta).cpu().numpy(), ref_map) < tol
            assert relerr(plan.adjoint(tm, 1.0).cpu().numpy(), ref_alm) < tol


def test_invalid_shape_is_rejected(gpu):
    from polyspec_amd._abi import ShtPlan, PolySpecHIPError
    plan = ShtPlan(16, 32, 0, 32)
    with pytest.raises(PolySpecHIPError):
        plan.set_shapes(2, [3, -1, -1, -1])      # 3 columns per block: no such kernel
    with pytest.raises(PolySpecHIPError):
        plan.set_shapes(2, [-1, -1, -1, 99])     # no adjoint shape 99
    with pytest.raises(PolySpecHIPError):
        plan.shape_candidates(2, 7)


@pytest.mark.parametrize("nside,lmax,prec,spin,nmaps", [
    (128, 256, 32, 0, 8), (128, 256, 32, 2, 4), (64, 128, 64, 0, 1)])
def test_autotuned_shapes_match_oracle(gpu, nside, lmax, prec, spin, nmaps):
    from polyspec_amd._abi import ShtPlan
    print(f"autotune nside={nside} lmax={lmax} prec={prec} spin={spin} nmaps={nmaps}", flush=True)
    a, m, ref_map, ref_alm = _case(nside, lmax, spin, nmaps, 5000 + nside + nmaps + 100 * spin)
    plan = ShtPlan(nside, lmax, spin, prec, autotune=True)
    times = plan.autotune(nmaps)
    shapes = plan.shapes(nmaps)
    for stage, name in enumerate(ShtPlan.SHAPE_STAGES):
        kept, builtin = times[name]
        assert kept >= 0 and builtin >= 0 and kept <= builtin
        assert shapes[stage] in plan.shape_candidates(nmaps, stage)
    tol = 1e-11 if prec == 64 else 2e-6
    for _ in range(2):  # the tuned shapes stay in place over calls
        out = plan.synthesis(torch.from_numpy(a).to(gpu, plan.cdtype)).cpu().numpy()
        assert relerr(out, ref_map) < tol
        out = plan.adjoint(torch.from_numpy(m).to(gpu, plan.rdtype), 1.0).cpu().numpy()
        assert relerr(out, ref_alm) < tol
    assert plan.shapes(nmaps) == shapes
