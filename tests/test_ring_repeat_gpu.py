"""Ring transforms launched many times in one process give bit-identical results.

Every ring-synthesis launch shape (unstaged radix-2^2 and self-sorting kernels, and the
LDS-staged kernels) and the ring adjoint are repeated on the same input: a read of LDS that a
missing barrier leaves unordered against another wave's write (as the Bluestein zero padding and
the fold once were) makes the output depend on wave timing, which repeats expose.  Numerical
agreement with the FP64 oracle is covered by tests/test_sht_tune_gpu.py; here the reference is
the first launch.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _alm_t(plan, ncol, seed):
    lmax = plan.lmax
    rng = np.random.default_rng(seed)
    ld = -(-ncol // plan.lmajor_multiple) * plan.lmajor_multiple
    l = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    a = (rng.standard_normal((plan.nlm, ld)) + 1j * rng.standard_normal((plan.nlm, ld))) / np.maximum(l, 1)[:, None]
    a[:, ncol:] = 0
    return a


@pytest.mark.parametrize("nside,ncol,prec,reps", [(512, 80, 32, 20), (128, 17, 32, 40), (256, 6, 64, 20),
                                                  (1024, 6, 32, 10), (1024, 4, 64, 5)])
def test_ring_synthesis_repeats_are_bit_identical(gpu, nside, ncol, prec, reps):
    """Also: every shape of a ring class gives the maps of every other shape to rounding (the
    staged, unstaged, self-sorting and radix-2^2 kernels are independent implementations; at
    Nside 1024 the Bluestein rings take 8192-point transforms)."""
    from polyspec_amd._abi import ShtPlan
    plan = ShtPlan(nside, 2 * nside, 0, prec)
    alm_t = torch.from_numpy(_alm_t(plan, ncol, 17 + nside + ncol)).to(gpu, plan.cdtype)
    phase = plan.phase_buffer(ncol)
    plan.legendre_synthesis_lmajor(alm_t, ncol, phase)
    out = torch.empty((ncol, 1, plan.npix), dtype=plan.rdtype, device=gpu)
    tol = 2e-6 if prec == 32 else 1e-12
    first = None
    for stage in ("ring_pow2", "ring_bluestein"):
        for c in plan.shape_candidates(ncol, stage):
            print(f"nside={nside} ncol={ncol} prec={prec} stage={stage} shape={c}", flush=True)
            shapes = [-1] * 4
            shapes[plan.SHAPE_STAGES.index(stage)] = c
            plan.set_shapes(ncol, shapes)
            ref = plan.ring_synthesis(phase, ncol).clone()
            assert torch.isfinite(ref).all()
            if first is None:
                first = ref
            err = float((ref.double() - first.double()).abs().max() / first.double().abs().max())
            assert err < tol, (stage, c, err)
            for _ in range(reps):
                plan.ring_synthesis(phase, ncol, out=out)
                assert torch.equal(out, ref), "ring synthesis shape %d differs between launches" % c


@pytest.mark.parametrize("nside,nmaps,spin,prec", [(512, 16, 0, 32), (256, 4, 2, 32), (128, 3, 0, 64)])
def test_adjoint_repeats_are_bit_identical(gpu, nside, nmaps, spin, prec):
    from polyspec_amd._abi import ShtPlan
    plan = ShtPlan(nside, 2 * nside, spin, prec)
    rng = np.random.default_rng(5 + nside + spin)
    maps = torch.from_numpy(rng.standard_normal((nmaps, plan.ncomp, plan.npix))).to(gpu, plan.rdtype)
    ref = plan.adjoint(maps, 1.0).clone()
    assert torch.isfinite(torch.view_as_real(ref)).all()
    out = torch.empty_like(ref)
    for _ in range(20):
        plan.adjoint(maps, 1.0, out=out)
        assert torch.equal(out, ref)
