"""GPU parity of the gNL-loc / tauNL-loc estimator against the FP64 oracle.

The oracle runs the reference's per-simulation loops; the GPU path runs either
the same loops (batch_sims=False) or the collapsed Monte-Carlo sums
(batch_sims=True).  Same inputs on both sides: data map, simulation alm, Fisher
GRFs and radial tables.

Tolerances: FP64 kernels 1e-8 of |t4| (numerator) / of max|F| (Fisher);
FP32 kernels 2e-4 (the amplitude target is 1e-4 sigma).
"""
import numpy as np
import pytest
import torch

import oracle
from oracle.tspec import OracleTSpec
from estimator_setup import grf_alm, radial_nodes, spectra, transfer

pytestmark = pytest.mark.gpu


def build(nside, lmax, pol, nr, templates, precision, batch_sims, nk=60, mask=None):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    raw, tot, beam = spectra(nside, pol)
    base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=precision)
    mask = np.ones(12 * nside ** 2) if mask is None else mask
    W = Weightings(base, mask, dict(raw))
    k, Tl = transfer(lmax, pol, nk)
    r, w = radial_nodes(nr)
    tm = TSpecTemplate(base, mask, W.applySinv_ideal, templates, k, Tl, 2, lmax, r_values=r,
                       r_weights={t: w for t in templates}, batch_sims=batch_sims)
    orc = OracleTSpec(nside, 2, lmax, tm.Lmin, tm.Lmax, tot, base.beam, pol, templates, tm.plLXs[0], tm.qlXs,
                      tm.FLLs[0] if 'tauNL-loc' in templates else None, tm.r_weights, mask=mask)
    return base, tm, orc, tot


CASES = [
    # nside, lmax, pol, nr, templates, precision, batch
    (16, 30, False, 4, ['gNL-loc', 'tauNL-loc'], 64, True),
    (16, 30, False, 4, ['gNL-loc', 'tauNL-loc'], 64, False),
    (16, 24, True, 3, ['gNL-loc', 'tauNL-loc'], 64, True),
    (16, 30, False, 4, ['gNL-loc', 'tauNL-loc'], 32, True),
    (16, 24, True, 3, ['gNL-loc'], 32, True),
]


@pytest.mark.parametrize("nside,lmax,pol,nr,templates,precision,batch", CASES)
def test_numerator_matches_oracle(gpu, nside, lmax, pol, nr, templates, precision, batch):
    print(f"numerator nside={nside} lmax={lmax} pol={pol} nr={nr} prec={precision} batch={batch}", flush=True)
    base, tm, orc, tot = build(nside, lmax, pol, nr, templates, precision, batch)
    rng = np.random.default_rng(42)
    L = 3 * nside - 1
    data_alm = grf_alm(rng, tot, L, pol)
    data = oracle.to_map(data_alm, nside, L)
    sims = [(grf_alm(rng, tot, lmax, pol), grf_alm(rng, tot, lmax, pol)) for _ in range(2)]
    tm.load_sims(lambda ii: sims[ii], len(sims), input_type='harmonic')
    got = tm.Tl_numerator(data)
    ref = orc.numerator(data, sims, input_type='map', sim_type='harmonic')
    ref4 = orc.numerator(data, sims, input_type='map', sim_type='harmonic', include_disconnected_term=False)
    tol = 1e-8 if precision == 64 else 2e-4
    for i, t in enumerate(templates):
        assert abs(got[i] - ref[i]) <= tol * abs(ref4[i]), (t, got[i], ref[i], ref4[i])
    # a second numerator reuses the cached t0 (tspec_template.py:1361-1365)
    got2 = tm.Tl_numerator(data)
    assert np.allclose(got2, got, rtol=1e-12 if precision == 64 else 1e-6, atol=0)


@pytest.mark.parametrize("precision,batch", [(64, False), (32, True)])
def test_numerator_at_the_reference_cpu_workload_size(gpu, precision, batch):
    """The reference's CPU-runnable gNL-loc workload at its own size: T only, unit mask,
    Nside 64, lmin 2, lmax 128, 10 radial nodes, transfer on 200 log-spaced k; the FP64
    per-simulation loop and the batched FP32 path against the FP64 oracle."""
    print(f"numerator nside=64 lmax=128 nr=10 prec={precision} batch={batch}", flush=True)
    nside, lmax, nr, templates = 64, 128, 10, ['gNL-loc']
    base, tm, orc, tot = build(nside, lmax, False, nr, templates, precision, batch, nk=200)
    rng = np.random.default_rng(64)
    L = 3 * nside - 1
    data = oracle.to_map(grf_alm(rng, tot, L, False), nside, L)
    sims = [(grf_alm(rng, tot, lmax, False), grf_alm(rng, tot, lmax, False)) for _ in range(2)]
    tm.load_sims(lambda ii: sims[ii], len(sims), input_type='harmonic')
    got = tm.Tl_numerator(data)
    ref = orc.numerator(data, sims, input_type='map', sim_type='harmonic')
    ref4 = orc.numerator(data, sims, input_type='map', sim_type='harmonic', include_disconnected_term=False)
    tol = 1e-8 if precision == 64 else 2e-4
    assert abs(got[0] - ref[0]) <= tol * abs(ref4[0]), (got[0], ref[0], ref4[0])


@pytest.mark.parametrize("nside,lmax,pol,nr,templates,precision", [
    (16, 30, False, 4, ['gNL-loc', 'tauNL-loc'], 64), (16, 24, True, 3, ['gNL-loc', 'tauNL-loc'], 64),
    (16, 30, False, 4, ['gNL-loc', 'tauNL-loc'], 32)])
def test_fisher_matches_oracle(gpu, nside, lmax, pol, nr, templates, precision):
    print(f"fisher nside={nside} lmax={lmax} pol={pol} nr={nr} prec={precision}", flush=True)
    base, tm, orc, tot = build(nside, lmax, pol, nr, templates, precision, True)
    rng = np.random.default_rng(7)
    # GRFs with the beam-deconvolved fiducial spectrum (generate_data(deconvolve_beam=True))
    a_maps = []
    for _ in range(2):
        a = grf_alm(rng, tot, lmax, pol)
        b = base.beam[:, oracle.getlm(lmax)[0]]
        a_maps.append(np.where(b != 0, a / np.where(b != 0, b, 1), 0))
    got = tm.fisher_contribution_from_grfs(a_maps)
    ref = orc.fisher_from_grfs(a_maps)
    tol = 1e-8 if precision == 64 else 2e-4
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (got, ref)
    got_d = tm.fisher_contribution_from_grfs([torch.from_numpy(a).to(gpu) for a in a_maps])
    assert np.abs(got_d - ref).max() <= tol * np.abs(ref).max(), (got_d, ref)


@pytest.mark.parametrize("pol,precision", [(True, 32), (True, 64), (False, 32)])
def test_graph_numerator_matches_eager(gpu, pol, precision):
    """capture_graph=True: the numerator's device work replayed from a HIP graph equals the eager
    numerator on every replay (three data maps, each twice, eager work in between)."""
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    print(f"graph numerator pol={pol} prec={precision}", flush=True)
    nside, lmax, nr = 16, 24, 3
    templates = ['gNL-loc', 'tauNL-loc']
    raw, tot, beam = spectra(nside, pol)
    rng = np.random.default_rng(31)
    L = 3 * nside - 1
    sims = [(grf_alm(rng, tot, lmax, pol), grf_alm(rng, tot, lmax, pol)) for _ in range(2)]
    k, Tl = transfer(lmax, pol)
    r, w = radial_nodes(nr)
    tms = []
    for graph in (False, True):
        # fixed (built-in) launch shapes: two separately tuned plans could pick different shapes,
        # whose FP32 results differ in rounding
        base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=precision, autotune=False)
        mask = np.ones(12 * nside ** 2)
        W = Weightings(base, mask, dict(raw))
        tm = TSpecTemplate(base, mask, W.applySinv_ideal, templates, k, Tl, 2, lmax, r_values=r,
                           r_weights={t: w for t in templates}, batch_sims=True, capture_graph=graph)
        tm.load_sims(lambda ii: sims[ii], len(sims), input_type='harmonic')
        tms.append(tm)
    rdt = torch.float32 if precision == 32 else torch.float64
    datas = [torch.from_numpy(oracle.to_map(grf_alm(rng, tot, L, pol), nside, L)).to(gpu, rdt).reshape(-1, 12 * nside ** 2)
             for _ in range(3)]
    for rep in range(2):
        for d in datas:
            want = tms[0].Tl_numerator(d)
            got = tms[1].Tl_numerator(d)
            assert np.allclose(got, want, rtol=1e-12, atol=0), (rep, got, want)
    assert tms[1]._graph is not None


@pytest.mark.parametrize("pol,precision", [(False, 64), (True, 64), (True, 32)])
def test_masked_sky_numerator_and_fisher_match_oracle(gpu, pol, precision):
    """The masked-sky branches: masked data and simulation maps through the map-space S^-1, and
    the Fisher contribution's 'Sinv' GRFs S^-1(mask to_map(beam a)), 'Ainv' GRFs applyAinv at
    base lmax, and the 'Ainv' weighting of the Q derivatives S^-1(mask to_map(beam Q))
    (tspec_template.py:2556-2572, 2616-2644), with a cut-plus-holes apodised mask."""
    from polyspec_amd.masks import galactic_cut_mask
    nside, lmax, nr = 16, 24, 3
    templates = ['gNL-loc', 'tauNL-loc']
    print(f"masked numerator+fisher nside={nside} lmax={lmax} pol={pol} prec={precision}", flush=True)
    mask = galactic_cut_mask(nside, cut_deg=20., nholes=12, hole_deg=6., apo_deg=8., seed=5)
    base, tm, orc, tot = build(nside, lmax, pol, nr, templates, precision, True, mask=mask)
    assert not tm.ones_mask and orc.mask is not None
    rng = np.random.default_rng(99)
    L = 3 * nside - 1
    masked = lambda: mask * oracle.to_map(grf_alm(rng, tot, L, pol), nside, L)
    data = masked()
    sims = [(masked(), masked()) for _ in range(2)]
    tm.load_sims(lambda ii: sims[ii], len(sims), input_type='map')
    got = tm.Tl_numerator(data)
    ref = orc.numerator(data, sims, input_type='map', sim_type='map')
    ref4 = orc.numerator(data, sims, input_type='map', sim_type='map', include_disconnected_term=False)
    tol = 1e-8 if precision == 64 else 2e-4
    for i, t in enumerate(templates):
        assert abs(got[i] - ref[i]) <= tol * abs(ref4[i]), (t, got[i], ref[i], ref4[i])
    # Fisher GRFs at base lmax with the beam deconvolved (generate_data(deconvolve_beam=True))
    ls = oracle.getlm(L)[0]
    a_maps = []
    for _ in range(2):
        a = grf_alm(rng, tot, L, pol)
        b = base.beam[:, ls]
        a_maps.append(np.where(b != 0, a / np.where(b != 0, b, 1), 0))
    got = tm.fisher_contribution_from_grfs(a_maps)
    ref = orc.fisher_from_grfs(a_maps)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (got, ref)
    # the same GRFs handed over as device tensors (generate_data with grf_device=True)
    got_d = tm.fisher_contribution_from_grfs([torch.from_numpy(a).to(gpu) for a in a_maps])
    assert np.abs(got_d - ref).max() <= tol * np.abs(ref).max(), (got_d, ref)
