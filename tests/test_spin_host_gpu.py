"""The host-side +-s recombination of PolySpec (backend 'hip') against brute-force sums of
spin-weighted harmonics at every HEALPix pixel centre.

Reference methods: to_map_spin / to_lm_spin / compute_spin_transform_map (base.py:234-297),
to_map_vec(output_spin > 0) / to_lm_vec(spin > 0) (base.py:178-232) with to_plus_minus,
to_plus_minus_complex, to_real_imag (ideal_fisher.pyx:237-322).

What they must equal (HEALPix conventions, every m including m < 0):
  compute_spin_transform_map(a, s)  = sum_lm a_lm sY_lm, a_{l,-m} = (-1)^m conj(a_lm), s = +-1, +-2, +-3
  to_map_spin(lp, lm, s)[0]         = sum_{m>=0} lp_lm sY_lm + sum_{m>0} (-1)^(m+s) conj(lm_lm) sY_{l,-m};
                                      [1] is its complex conjugate
  to_lm_spin(P, conj P, s)          = dOmega sum_p P(p) conj(sY_lm(p)),  dOmega sum_p conj P(p) conj(-sY_lm(p))
  to_map_vec(a, output_spin=s)      = [compute_spin_transform_map(a, s), its conjugate]
  to_lm_vec([P, conj P], spin=s)    = to_lm_spin(P, conj P, s) (the Cython to_real_imag, with its /2)
sY_lm is the Goldberg et al. (1967) closed form (tests/golden/make_golden.py), itself pinned
against eth-derivatives of scipy's Y_lm in tests/test_oracle_sht.py.
"""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import getlm, goldberg_sylm, healpix_centres  # noqa: E402

pytestmark = pytest.mark.gpu

NSIDE, LMAX = 4, 8


def _Y(s, th, ph, ls, ms):
    """(sY_lm, sY_{l,-m}) for the m >= 0 index set, (nlm, npix) each."""
    pos = np.array([goldberg_sylm(s, l, m, th, ph) for l, m in zip(ls, ms)])
    neg = np.array([goldberg_sylm(s, l, -m, th, ph) for l, m in zip(ls, ms)])
    return pos, neg


def _base(precision):
    from polyspec_amd.base import PolySpec
    l = np.arange(3 * NSIDE)
    cl = np.ones_like(l, dtype=float)
    return PolySpec(NSIDE, {'TT': cl}, beam=[np.ones_like(cl)], pol=False, precision=precision)


def _rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("s", [1, 2, 3])
def test_spin_recombination_vs_bruteforce(gpu, s, precision):
    print(f"spin host recombination s={s} precision={precision}", flush=True)
    base = _base(precision)
    tol = 1e-12 if precision == 64 else 3e-6
    th, ph = healpix_centres(NSIDE)
    ls, ms = getlm(LMAX)
    pos = ms > 0
    sg = (-1.) ** s
    Yp, Yn = _Y(s, th, ph, ls, ms)
    Ymp, Ymn = _Y(-s, th, ph, ls, ms)
    rng = np.random.default_rng(10 * s + precision)

    # compute_spin_transform_map: sum over every m of a scalar a_lm times +-sY_lm
    a = rng.standard_normal(len(ls)) + 1j * rng.standard_normal(len(ls))
    a[ms == 0] = a[ms == 0].real
    a[ls < s] = 0
    conj_neg = (-1.) ** ms[pos] * np.conj(a[pos])
    for sign, (P_, N_) in ((1, (Yp, Yn)), (-1, (Ymp, Ymn))):
        brute = a @ P_ + conj_neg @ N_[pos]
        got = base.compute_spin_transform_map([a], sign * s, lmax=LMAX)[0]
        assert _rel(got, brute) < tol, (sign, _rel(got, brute))

    # to_map_spin with independent +s / -s harmonics (m = 0 rows of a real spin field)
    lp = rng.standard_normal(len(ls)) + 1j * rng.standard_normal(len(ls))
    lmn = rng.standard_normal(len(ls)) + 1j * rng.standard_normal(len(ls))
    lp[ls < s] = 0
    lmn[ls < s] = 0
    lmn[ms == 0] = sg * np.conj(lp[ms == 0])
    brute = lp @ Yp + ((-1.) ** (ms[pos] + s) * np.conj(lmn[pos])) @ Yn[pos]
    mp, mm = base.to_map_spin(lp, lmn, s, lmax=LMAX)
    assert _rel(mp, brute) < tol
    assert _rel(mm, np.conj(brute)) < tol

    # to_lm_spin of a spin field and its conjugate
    npix = 12 * NSIDE ** 2
    P = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    dA = 4 * np.pi / npix
    want_p, want_m = dA * (np.conj(Yp) @ P), dA * (np.conj(Ymp) @ np.conj(P))
    got_p, got_m = base.to_lm_spin(P, np.conj(P), s, lmax=LMAX)
    assert _rel(got_p, want_p) < tol and _rel(got_m, want_m) < tol

    # batched forms: two maps / two alm sets per call
    A = np.stack([a, 0.5j * a * (ls >= s)])
    A[1, ms == 0] = 0
    vec = base.to_map_vec(A, output_spin=s, lmax=LMAX)
    for b in range(2):
        cn = (-1.) ** ms[pos] * np.conj(A[b, pos])
        brute = A[b] @ Yp + cn @ Yn[pos]
        assert _rel(vec[0][b], brute) < tol and _rel(vec[1][b], np.conj(brute)) < tol
    Pm = np.stack([P, P[::-1] * 0.3])
    lmv = base.to_lm_vec([Pm, np.conj(Pm)], spin=s, lmax=LMAX)
    for b in range(2):
        wp, wm = dA * (np.conj(Yp) @ Pm[b]), dA * (np.conj(Ymp) @ np.conj(Pm[b]))
        assert _rel(lmv[0][b], wp) < tol and _rel(lmv[1][b], wm) < tol
