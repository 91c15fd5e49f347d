"""Uninitialised-memory check of the HIP paths.

Every case runs twice: once after torch's caching allocator and the library's cached
workspace have been refilled with NaN, once after they have been refilled with zeros.
The kernels are deterministic for fixed launch shapes (the built-in ones here: two
autotuning passes may keep different shapes, whose FP32 sums round differently), so the
two results must be finite and bit-identical; a
kernel that reads device memory it never wrote fails here even where fresh device memory
happens to hold zeros (or the values a previous identical run left behind).
"""
import numpy as np
import pytest
import torch

import oracle
from estimator_setup import grf_alm, radial_nodes, spectra, transfer

pytestmark = pytest.mark.gpu


def _refill(value):
    """Return 2 GiB of large blocks and 128 MiB of small blocks, filled with value, to torch's cache."""
    from polyspec_amd import _abi
    torch.cuda.synchronize()
    _abi._workspaces.clear()
    torch.cuda.empty_cache()
    big = [torch.full((1 << 27,), value, dtype=torch.float32, device="cuda") for _ in range(4)]
    small = [torch.full((1 << 17,), value, dtype=torch.float32, device="cuda") for _ in range(256)]
    torch.cuda.synchronize()
    del big, small


def _twice(fn):
    out = []
    for v in (float("nan"), 0.0):
        _refill(v)
        r = fn()
        torch.cuda.synchronize()
        out.append([np.asarray(x) for x in r])
    for a, b in zip(*out):
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b)


@pytest.mark.parametrize("nside,lmax,spin,prec,nmap", [
    (32, 64, 0, 32, 1), (32, 64, 0, 32, 3), (32, 64, 0, 32, 20), (32, 64, 0, 64, 1), (32, 64, 0, 64, 5),
    (32, 64, 1, 32, 1), (32, 64, 1, 32, 3), (32, 64, 1, 32, 21), (32, 64, 2, 32, 1), (32, 64, 2, 32, 9),
    (32, 64, 2, 64, 1), (32, 64, 2, 64, 3), (64, 191, 2, 32, 2), (64, 128, 0, 32, 33)])
def test_sht_reads_only_written_memory(gpu, nside, lmax, spin, prec, nmap):
    from polyspec_amd._abi import ShtPlan
    print(f"poison sht nside={nside} lmax={lmax} spin={spin} prec={prec} nmap={nmap}", flush=True)
    rng = np.random.default_rng(5)
    ls, ms = oracle.getlm(lmax)
    nc = 2 if spin else 1
    a = rng.standard_normal((nmap, nc, len(ls))) + 1j * rng.standard_normal((nmap, nc, len(ls)))
    a[..., ms == 0] = a[..., ms == 0].real
    a[..., ls < spin] = 0
    m = rng.standard_normal((nmap, nc, 12 * nside ** 2))

    def run():
        plan = ShtPlan(nside, lmax, spin, prec)
        y = plan.synthesis(torch.from_numpy(a).to(gpu, plan.cdtype)).cpu()
        yt = plan.adjoint(torch.from_numpy(m).to(gpu, plan.rdtype), 0.5).cpu()
        return y.numpy(), yt.numpy()
    _twice(run)


@pytest.mark.parametrize("prec", [32, 64])
def test_bspec_reads_only_written_memory(gpu, prec):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.bspec_bin import BSpecBin
    from polyspec_amd.weighting import Weightings
    print(f"poison bspec prec={prec}", flush=True)
    nside = 16
    raw, tot, beam = spectra(nside, True)
    L = 3 * nside - 1
    rng = np.random.default_rng(8)
    maps = [oracle.to_map(grf_alm(rng, tot, L, True), nside, L) for _ in range(3)]

    def run():
        base = PolySpec(nside, dict(tot), beam=beam, pol=True, precision=prec, autotune=False)
        W = Weightings(base, np.ones(12 * nside ** 2), dict(raw))
        bs = BSpecBin(base, np.ones(12 * nside ** 2), W.applySinv_ideal, [2, 6, 12, 20, 30, 42],
                      fields=['TTT', 'TTE', 'TEE', 'EEE'], parity='both')
        bs.load_sims(lambda i: maps[1 + i], 2, preload=True)
        return (bs.Bl_numerator(maps[0]),)
    _twice(run)


@pytest.mark.parametrize("pol,prec", [(False, 32), (True, 32), (True, 64)])
def test_template_reads_only_written_memory(gpu, pol, prec):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    print(f"poison template pol={pol} prec={prec}", flush=True)
    nside, lmax, nr = 16, 24, 3
    templates = ['gNL-loc', 'tauNL-loc']
    raw, tot, beam = spectra(nside, pol)
    rng = np.random.default_rng(9)
    L = 3 * nside - 1
    data = oracle.to_map(grf_alm(rng, tot, L, pol), nside, L)
    sims = [(grf_alm(rng, tot, lmax, pol), grf_alm(rng, tot, lmax, pol)) for _ in range(2)]
    k, Tl = transfer(lmax, pol)
    r, w = radial_nodes(nr)

    def run():
        base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=prec, autotune=False)
        mask = np.ones(12 * nside ** 2)
        W = Weightings(base, mask, dict(raw))
        tm = TSpecTemplate(base, mask, W.applySinv_ideal, templates, k, Tl, 2, lmax, r_values=r,
                           r_weights={t: w for t in templates}, batch_sims=True)
        tm.load_sims(lambda ii: sims[ii], len(sims), input_type='harmonic')
        return (tm.Tl_numerator(data),)
    _twice(run)
