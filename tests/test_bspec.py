"""Binned bispectrum: the oracle's H maps against brute-force spin-weighted harmonics
(Goldberg et al. 1967, pinned against eth-derivatives of scipy's Y_lm in
test_oracle_sht.py), the negative-spin identity of that closed form, and the GPU
BSpecBin numerator against the oracle on the same seeded inputs."""
import os
import sys

import numpy as np
import pytest

import oracle
from oracle.bspec import OracleBSpec
from estimator_setup import spectra

sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
from make_golden import goldberg_sylm, healpix_centres  # noqa: E402


def brute_H(a, s, lmax, th, ph):
    """sum_{l, m=-l..l} at_lm sY_lm with at_{l,-m} = (-1)^m conj(a_lm) (a real-field's stored m >= 0)."""
    ls, ms = oracle.getlm(lmax)
    out = np.zeros(len(th), complex)
    for i, (l, m) in enumerate(zip(ls, ms)):
        if a[i] == 0:
            continue
        out += a[i] * goldberg_sylm(s, l, m, th, ph)
        if m > 0:
            out += (-1) ** m * np.conj(a[i]) * goldberg_sylm(s, l, -m, th, ph)
    return out


def test_negative_spin_identity():
    """-sY_lm = (-1)^(s+m) conj(sY_l,-m) for the closed form used by the brute force."""
    th = np.linspace(0.2, 2.9, 7)
    ph = np.linspace(0.1, 6.0, 7)
    for s, l, m in [(2, 3, 1), (2, 4, -2), (1, 2, 2), (1, 5, -3)]:
        lhs = goldberg_sylm(-s, l, m, th, ph)
        rhs = (-1) ** (s + m) * np.conj(goldberg_sylm(s, l, -m, th, ph))
        assert np.allclose(lhs, rhs, atol=1e-13)


def test_oracle_H_maps_match_bruteforce():
    nside, lmax = 4, 11
    th, ph = healpix_centres(nside)
    rng = np.random.default_rng(5)
    ls, ms = oracle.getlm(lmax)
    h = rng.standard_normal((1, len(ls))) + 1j * rng.standard_normal((1, len(ls)))
    h[:, ms == 0] = h[:, ms == 0].real
    o = OracleBSpec(nside, lmax, [2, 5, 9], ['TTT'], 'even', False)
    p1, m2 = o.H_maps(h)
    for b in range(2):
        a = (o.ell_bins[b] * h.conj())[0]
        assert np.abs(p1[b][0] - brute_H(a, 1, lmax, th, ph)).max() < 1e-12 * np.abs(p1[b][0]).max()
        assert np.abs(m2[b][0] - brute_H(a, -2, lmax, th, ph)).max() < 1e-12 * np.abs(m2[b][0]).max()


def test_oracle_bins_and_symmetry():
    o = OracleBSpec(8, 23, [2, 6, 12, 20], ['TTT', 'TTE', 'EEE'], 'both', True)
    # TTT: b1 <= b2 <= b3 with the centre triangle condition
    c = [4., 9., 16.]
    want = sum(1 for a in range(3) for b in range(a, 3) for d in range(b, 3)
               if abs(c[a] - c[b]) <= c[d] <= c[a] + c[b])
    assert sum(1 for x in o.all_bins if x[:3] == (0, 0, 0)) == want
    # fully symmetric TTT triplets carry 6, two equal indices 2
    for (u1, u2, u3, b1, b2, b3, _), s in zip(o.all_bins, o.sym):
        if (u1, u2, u3) == (0, 0, 0):
            assert s == (6 if b1 == b2 == b3 else 2 if (b1 == b2 or b2 == b3) else 1)
    assert len(o.sym) == 2 * len(o.all_bins)


def _setup(nside, pol, prec, nsims):
    import torch
    from polyspec_amd.base import PolySpec
    from polyspec_amd.bspec_bin import BSpecBin
    from polyspec_amd.weighting import Weightings
    raw, tot, beam = spectra(nside, pol)
    if not pol:
        tot = {'TT': tot['TT']}
        beam = [beam[0]]
    base = PolySpec(nside, dict(tot), beam=beam, pol=pol, precision=prec)
    mask = np.ones(12 * nside ** 2)
    W = Weightings(base, mask, dict(raw) if pol else {'TT': raw['TT']})
    L = base.lmax
    bins = [2, 6, 12, 20, 30, 42]
    fields = ['TTT', 'TTE', 'TEE', 'EEE'] if pol else ['TTT']
    bs = BSpecBin(base, mask, W.applySinv_ideal, bins, fields=fields, parity='both' if pol else 'even')
    rng = np.random.default_rng(17)

    def field(seed):
        r = np.random.default_rng(seed)
        from estimator_setup import grf_alm
        return oracle.to_map(grf_alm(r, tot, L, pol), nside, L)
    datas = [field(100 + i) for i in range(1 + nsims)]
    return base, W, bs, bins, fields, datas


@pytest.mark.gpu
@pytest.mark.parametrize("pol,prec,nsims", [(False, 64, 0), (True, 64, 2), (True, 32, 2)])
def test_gpu_bspec_numerator_matches_oracle(gpu, pol, prec, nsims):
    import torch
    print(f"bspec pol={pol} prec={prec} nsims={nsims}", flush=True)
    nside = 16
    base, W, bs, bins, fields, datas = _setup(nside, pol, prec, nsims)
    o = OracleBSpec(nside, base.lmax, bins, bs.fields, bs.parity, pol)
    assert [tuple(x) for x in bs.all_bins] == [tuple(x) for x in o.all_bins]
    hs = [W.applySinv_ideal(d) for d in datas]  # numpy in -> numpy out, FP64 host copy of S^-1 d
    if nsims:
        bs.load_sims(lambda i: datas[1 + i], nsims, preload=True)
    got = bs.Bl_numerator(datas[0], include_linear_term=bool(nsims))
    ref = o.numerator(np.asarray(hs[0]), [np.asarray(h) for h in hs[1:]])
    scale = np.abs(ref).max()
    tol = 1e-9 if prec == 64 else 2e-4
    assert np.abs(got - ref).max() < tol * scale, (np.abs(got - ref).max() / scale)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,nmap,npix,ntrip", [(64, 7, 5001, 300), (32, 7, 5001, 300), (32, 60, 20000, 4500),
                                                  (64, 128, 3000, 70)])
def test_gpu_triple_and_linear_sums_match_numpy(gpu, prec, nmap, npix, ntrip):
    """ps_bspec_triple_sums (three stacks and A = B) and ps_bspec_linear_sums against direct
    FP64 sums; ragged pixel tiles, more triplets than one chunk, the 128-map maximum."""
    import torch
    from polyspec_amd._abi import load, ptr, stream_handle, check
    print(f"bspec sums prec={prec} nmap={nmap} npix={npix} ntrip={ntrip}", flush=True)
    rng = np.random.default_rng(nmap + ntrip)
    X = [rng.standard_normal((nmap, npix)) + 1j * rng.standard_normal((nmap, npix)) for _ in range(4)]
    trip = rng.integers(0, nmap, (ntrip, 3)).astype(np.int32)
    cdt = torch.complex64 if prec == 32 else torch.complex128
    Xd = [torch.from_numpy(x).to(gpu, cdt).contiguous() for x in X]
    Xr = [x.astype(np.complex64).astype(np.complex128) if prec == 32 else x for x in X]
    td = torch.from_numpy(trip).to(gpu)
    lib = load()
    nb = int(lib.ps_bspec_scratch_bytes(ntrip))
    scr = torch.empty(nb, dtype=torch.uint8, device=gpu)

    def cyc(A, B, C):
        a, b, c = trip[:, 0], trip[:, 1], trip[:, 2]
        return (A[a] * B[b] * C[c] + A[b] * B[c] * C[a] + A[c] * B[a] * C[b]).sum(1)

    def run(fn, stacks, scale=1.0):
        out = torch.zeros((ntrip, 2), dtype=torch.float64, device=gpu)
        check(fn(prec, *[ptr(s) for s in stacks], nmap, npix, ptr(td), ntrip, scale, ptr(scr), nb, ptr(out), 0,
                 stream_handle()))
        o = out.cpu().numpy()
        return o[:, 0] + 1j * o[:, 1]
    tol = 1e-11 if prec == 64 else 2e-5
    P, Q, M, N = Xr
    want3 = cyc(P, Q, M)
    got3 = run(lib.ps_bspec_triple_sums, (Xd[0], Xd[1], Xd[2]), 1.0)
    assert np.abs(got3 - want3).max() < tol * np.abs(want3).max()
    wantc = cyc(P, P, M)
    gotc = run(lib.ps_bspec_triple_sums, (Xd[0], Xd[0], Xd[2]), 1.0)
    assert np.abs(gotc - wantc).max() < tol * np.abs(wantc).max()
    wantl = cyc(P, Q, N) + cyc(Q, P, N) + cyc(Q, Q, M)
    gotl = run(lib.ps_bspec_linear_sums, tuple(Xd), 0.5) / 0.5
    assert np.abs(gotl - wantl).max() < tol * np.abs(wantl).max()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [32, 64])
def test_gpu_multi_segment_sums_match_numpy(gpu, prec):
    """ps_bspec_sums_multi: segments of every mode, different stack sizes, pixel counts and
    triplet counts, two segments adding into one group, every launch shape, against direct FP64
    sums; the autotuner returns a valid shape."""
    import torch
    from polyspec_amd import bspec_bin as B
    from polyspec_amd._abi import load
    print(f"bspec multi prec={prec}", flush=True)
    rng = np.random.default_rng(23)
    cdt = torch.complex64 if prec == 32 else torch.complex128

    def stack(nmap, npix):
        x = rng.standard_normal((nmap, npix)) + 1j * rng.standard_normal((nmap, npix))
        xd = torch.from_numpy(x).to(gpu, cdt).contiguous()
        return xd, xd.cpu().to(torch.complex128).numpy()

    def cyc(A, Bm, C, trip):
        a, b, c = trip[:, 0], trip[:, 1], trip[:, 2]
        return (A[a] * Bm[b] * C[c] + A[b] * Bm[c] * C[a] + A[c] * Bm[a] * C[b]).sum(1)

    # group 0: cubic (7 maps, 5001 px) + triple; group 1: two linear segments (60 maps, 3000 px);
    # group 2: one linear segment of another size (12 maps, 777 px, 2100 triplets)
    specs = [(0, 1, 7, 5001, 300), (0, 0, 7, 5001, 300), (1, 2, 60, 3000, 4500), (1, 2, 60, 3000, 4500),
             (2, 2, 12, 777, 2100)]
    segs, want = [], {}
    trips = {}
    for g, mode, nmap, npix, ntrip in specs:
        if g not in trips:
            trips[g] = rng.integers(0, nmap, (ntrip, 3)).astype(np.int32)
        t = trips[g]
        td = torch.from_numpy(t).to(gpu)
        ns = {0: 3, 1: 2, 2: 4}[mode]
        st = [stack(nmap, npix) for _ in range(ns)]
        X = [x[1] for x in st]
        if mode == 0:
            w = cyc(X[0], X[1], X[2], t)
        elif mode == 1:
            w = cyc(X[0], X[0], X[1], t)
        else:
            P, Q, M, N = X
            w = cyc(P, Q, N, t) + cyc(Q, P, N, t) + cyc(Q, Q, M, t)
        want[g] = want.get(g, 0) + w
        segs.append((tuple(x[0] for x in st), td, g))
    tol = 1e-11 if prec == 64 else 2e-5
    nshape = load().ps_bspec_multi_shape_count(prec)
    assert nshape >= 4
    for shape in list(range(nshape)) + [-1]:
        outs = [torch.full((len(trips[g]), 2), np.nan, dtype=torch.float64, device=gpu) for g in range(3)]
        B.multi_sums(segs, outs, shape)
        for g in range(3):
            o = outs[g].cpu().numpy()
            got = o[:, 0] + 1j * o[:, 1]
            err = np.abs(got - want[g]).max() / np.abs(want[g]).max()
            assert err < tol, (shape, g, err)
    best, ms = B.multi_autotune(segs, 3)
    assert 0 <= best < nshape and ms[0] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [32, 64])
def test_gpu_bspec_fused_and_multi_resolution(gpu, prec):
    """Bl_numerator's fused launch equals the per-simulation launches, and Bl_numerator_multi
    over two resolutions equals each analysis's own numerator."""
    import torch
    from polyspec_amd.bspec_bin import Bl_numerator_multi
    print(f"bspec fused/multi prec={prec}", flush=True)
    runs = []
    for nside in (16, 32):
        base, W, bs, bins, fields, datas = _setup(nside, True, prec, 3)
        bs.load_sims(lambda i: datas[1 + i], 3, preload=True)
        bs.fused_sums = False
        sep = bs.Bl_numerator(datas[0])
        bs.fused_sums = True
        fused = bs.Bl_numerator(datas[0])
        tol = 1e-12 if prec == 64 else 1e-6
        assert np.abs(fused - sep).max() < tol * np.abs(sep).max()
        runs.append((bs, datas[0], fused))
    # one launch over both resolutions: other pixel ranges per segment, so FP32 tile sums round
    # differently
    tol = 1e-12 if prec == 64 else 1e-6
    multi = Bl_numerator_multi([r[0] for r in runs], [r[1] for r in runs])
    for (bs, d, fused), m in zip(runs, multi):
        assert np.abs(m - fused).max() < tol * np.abs(fused).max()
    cub = Bl_numerator_multi([r[0] for r in runs], [r[1] for r in runs], include_linear_term=False)
    for (bs, d, fused), c in zip(runs, cub):
        ref = bs.Bl_numerator(d, include_linear_term=False)
        assert np.abs(c - ref).max() < tol * np.abs(ref).max()
    # streamed simulations (preload=False): rounds of simulations whose H maps fit the budget,
    # each round one launch adding to the earlier ones.  A budget of one Nside-32 simulation's maps
    # gives one round of the three Nside-16 simulations (a quarter of that each), then one round
    # per Nside-32 simulation; a large budget, one round of all six.  The Nside-16 analysis streamed
    # beside a preloaded Nside-32 one covers the mixed case.
    keep = [(r[0].p1_H_maps, r[0].m2_H_maps) for r in runs]
    for r, (p1s, m2s) in zip(runs, keep):
        r[0].load_sims(lambda i, p1s=p1s, m2s=m2s: None, 3, preload=False)
        r[0].load_sim_data = lambda i, p1s=p1s, m2s=m2s: (p1s[i], m2s[i])
    per32 = 2 * keep[1][0][0].numel() * keep[1][0][0].element_size()
    for budget in (per32 / 2 ** 30, 1.0):
        streamed = Bl_numerator_multi([r[0] for r in runs], [r[1] for r in runs], sim_memory_gb=budget)
        for (bs, d, fused), m in zip(runs, streamed):
            assert np.abs(m - fused).max() < tol * np.abs(fused).max(), budget
    runs[1][0].preload = True
    runs[1][0].p1_H_maps, runs[1][0].m2_H_maps = keep[1]
    mixed = Bl_numerator_multi([r[0] for r in runs], [r[1] for r in runs], sim_memory_gb=per32 / 2 ** 30)
    for (bs, d, fused), m in zip(runs, mixed):
        assert np.abs(m - fused).max() < tol * np.abs(fused).max()


def test_stream_rounds_plan():
    """Rounds of streamed simulations: round-robin over analyses while the round's H maps fit the
    budget (at least one simulation per round), each analysis's share of a round loaded in one
    call, in simulation order."""
    import torch
    from polyspec_amd.bspec_bin import _stream_rounds

    class Fake:
        def __init__(self, n, tag):
            self.N_it, self.tag, self._trip, self.calls = n, tag, tag + '_trip', []

        def _load_sims_data(self, idx):
            self.calls.append(list(idx))
            return [((self.tag, i), (self.tag, -i)) for i in idx]

    a, b = Fake(5, 'a'), Fake(2, 'b')
    pa = torch.zeros(10, dtype=torch.float32)  # 80 bytes of H maps per simulation (two stacks)
    pb = torch.zeros(20, dtype=torch.float32)  # 160 bytes
    rounds = list(_stream_rounds([(a, pa, pa, 3), (b, pb, pb, 7)], 400))
    # round 1: a0 (80), b0 (240), a1 (320), b1 would be 480, a2 (400); round 2: a3, b1, a4
    assert a.calls == [[0, 1, 2], [3, 4]] and b.calls == [[0], [1]]
    assert [[(s[0][1], s[2]) for s in r] for r in rounds] == [
        [(('a', 0), 3), (('a', 1), 3), (('a', 2), 3), (('b', 0), 7)],
        [(('a', 3), 3), (('a', 4), 3), (('b', 1), 7)]]
    for r in rounds:
        for st, trip, g in r:
            tag = st[1][0]
            assert st[0] is (pa if tag == 'a' else pb) and st[3] == (tag, -st[1][1]) and trip == tag + '_trip'
    # a budget below one simulation: one simulation per round
    c = Fake(3, 'c')
    assert [len(r) for r in _stream_rounds([(c, pa, pa, 0)], 10)] == [1, 1, 1]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [64, 32])
def test_gpu_streamed_sims_batched_sinv(gpu, prec):
    """Streamed simulations filtered by the batched ideal S^-1 (Weightings.applySinv_ideal_batch,
    chunks of sinv_batch simulations in a round) against the same simulations preloaded, and the
    batched filter against one applySinv_ideal call per map set."""
    import torch
    from polyspec_amd.bspec_bin import Bl_numerator_multi
    print(f"bspec streamed batched S^-1 prec={prec}", flush=True)
    base, W, bs, bins, fields, datas = _setup(32, True, prec, 5)
    sims = [torch.as_tensor(d).to(base.device, base.rdtype).contiguous() for d in datas[1:]]
    # FP32: the batched and the single-map adjoints are different kernels (matrix-core and vector)
    # whose rounding differs by up to about 1e-5 of the largest filtered coefficient (C_tot^-1
    # raises the high-l ones); a wrong simulation, field or index differs at order one
    tol = 1e-12 if prec == 64 else 5e-5
    hb = W.applySinv_ideal_batch(torch.stack(sims))
    assert hb.shape == (5, 3, base.plan(base.lmax, 0).nlm)
    for j, m in enumerate(sims):
        h1 = W.applySinv_ideal(m)
        assert float((hb[j] - h1).abs().max()) <= tol * float(h1.abs().max()), j
    bs.load_sims(lambda i: sims[i], 5, preload=True)
    ref = Bl_numerator_multi([bs], [datas[0]])[0]
    bs.load_sims(lambda i: sims[i], 5, preload=False)
    assert bs._sinv_batch() is not None
    bs.sinv_batch = 2  # one round of five: batches of 2, 2 and a single simulation
    got = Bl_numerator_multi([bs], [datas[0]], sim_memory_gb=1.0)[0]
    assert np.abs(got - ref).max() < tol * np.abs(ref).max()
    # a budget below one simulation's maps: one simulation per round, the single-simulation path
    one = Bl_numerator_multi([bs], [datas[0]], sim_memory_gb=1e-9)[0]
    assert np.abs(one - ref).max() < tol * np.abs(ref).max()
