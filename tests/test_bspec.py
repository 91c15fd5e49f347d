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
