"""Gradient-only spin-s synthesis (ps_sht_synthesis_grad): the synthesis of (G, C = 0) from G alone,
against the full synthesis of (G, 0) on the same plan (FP32 matrix-core kernel of the (G, C)
columns, itself pinned to the FP64 oracle in test_sht_tune_gpu.py) and against the FP64 oracle.

The gradient-only kernel takes 16 maps per column tile instead of 8 and forms the I component
from the lambda- accumulator of the partner lane; a swapped sign or partner, or a wrong parity of
the south ring, moves the result by O(1) of the maps.  Map counts cover a partial tile, one and
several tiles, and a plan with ring slices and Bluestein rings.
"""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nside,lmax,spin,nmaps", [(16, 40, 1, 1), (32, 64, 2, 5), (64, 128, 1, 17),
                                                   (64, 150, 3, 33), (256, 512, 2, 20)])
def test_gradient_only_synthesis(gpu, nside, lmax, spin, nmaps):
    from polyspec_amd._abi import ShtPlan
    print(f"synthesis_grad nside={nside} lmax={lmax} spin={spin} nmaps={nmaps}", flush=True)
    rng = np.random.default_rng(100 + nside + spin + nmaps)
    ls, ms = oracle.getlm(lmax)
    G = (rng.standard_normal((nmaps, len(ls))) + 1j * rng.standard_normal((nmaps, len(ls)))) / np.maximum(ls, 1)
    G[:, ms == 0] = G[:, ms == 0].real
    G[:, ls < spin] = 0
    plan = ShtPlan(nside, lmax, spin, 32)
    Gd = torch.from_numpy(G).to(gpu, torch.complex64)
    got = plan.synthesis_grad(Gd).double().cpu().numpy()
    full = plan.synthesis(torch.stack([Gd, torch.zeros_like(Gd)], dim=1).contiguous()).double().cpu().numpy()
    scale = np.abs(full).max()
    assert np.abs(got - full).max() <= 2e-6 * scale
    if nside <= 64:
        GC = np.stack([G, np.zeros_like(G)], axis=1)
        ref = oracle.synthesis(GC, nside, lmax, spin)
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
    # repeated calls are bit-identical; an FP64 plan falls back to the full synthesis of (G, 0)
    again = plan.synthesis_grad(Gd).double().cpu().numpy()
    assert np.array_equal(again, got)
    # every Legendre shape of the batch (the one-tile kernel, the wide form at 2, 3, 4 tiles):
    # the same products in the same order per output, so the same maps
    for sh in plan.shape_candidates(nmaps, 'synthesis'):
        if sh == 0 or sh >= 65:
            plan.set_shapes(nmaps, [sh, -1, -1, -1])
            alt = plan.synthesis_grad(Gd).double().cpu().numpy()
            assert np.abs(alt - got).max() <= 2e-6 * scale, sh
    if nside <= 32:
        p64 = ShtPlan(nside, lmax, spin, 64)
        g64 = p64.synthesis_grad(torch.from_numpy(G).to(gpu)).cpu().numpy()
        assert np.abs(g64 - oracle.synthesis(np.stack([G, 0 * G], axis=1), nside, lmax, spin)).max() <= 1e-11 * scale
