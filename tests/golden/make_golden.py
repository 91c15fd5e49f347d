"""Generate the brute-force SHT fixtures in tests/golden/.

Spin 0: direct sums of scipy.special.sph_harm_y (Condon-Shortley phase, the
healpy/ducc convention) at every HEALPix pixel centre.  Spin s > 0: direct sums
of the Goldberg et al. (1967) closed form of sY_lm (``goldberg_sylm`` below),
with the HEALPix gradient/curl convention R + iI = -sum (G + iC) sY_lm
(polyspec/base.py:247-256, 276-286 describe the +-s recombination around it).
Large-l spot values of Y_lm come from mpmath's Ferrers function at 60 digits.

Pixel centres use the HEALPix RING geometry written out independently here
(Gorski et al. 2005 sec. 4), not the oracle's, so the fixture also pins the
oracle's geometry.

Run:  python tests/golden/make_golden.py   (writes sht_bruteforce.npz, ylm_spot.npz)
"""
import math
import os

import numpy as np
from scipy.special import sph_harm_y

HERE = os.path.dirname(os.path.abspath(__file__))


def healpix_centres(nside):
    th, ph = [], []
    for i in range(1, 4 * nside):
        ir = min(i, 4 * nside - i)
        if ir < nside:
            z = 1.0 - ir * ir / (3.0 * nside * nside)
            n = 4 * ir
            phi = np.pi / (2 * ir) * (np.arange(1, n + 1) - 0.5)
        else:
            z = (4.0 / 3.0) - 2.0 * ir / (3.0 * nside)
            n = 4 * nside
            s = (ir - nside + 1) % 2
            phi = np.pi / (2 * nside) * (np.arange(1, n + 1) - 1 + s / 2.0)
        if i > 2 * nside:
            z = -z
        th.append(np.full(n, math.acos(z)))
        ph.append(phi)
    return np.concatenate(th), np.concatenate(ph)


def getlm(lmax):
    ls, ms = [], []
    for m in range(lmax + 1):
        ls += list(range(m, lmax + 1))
        ms += [m] * (lmax + 1 - m)
    return np.array(ls), np.array(ms)


def goldberg_sylm(s, l, m, theta, phi):
    """Goldberg et al. (1967) spin-weighted harmonic sY_lm(theta, phi)."""
    s, l, m = int(s), int(l), int(m)
    if l < abs(s) or l < abs(m):
        return np.zeros_like(theta, dtype=complex)
    pref = (-1.0) ** m * math.sqrt(math.factorial(l + m) * math.factorial(l - m) * (2 * l + 1) /
                                   (4 * math.pi * math.factorial(l + s) * math.factorial(l - s)))
    sh, ch = np.sin(theta / 2), np.cos(theta / 2)
    tot = np.zeros_like(theta, dtype=complex)
    for r in range(0, l - s + 1):
        k = r + s - m
        if k < 0 or k > l + s:
            continue
        c = math.comb(l - s, r) * math.comb(l + s, k) * (-1.0) ** (l - r - s)
        # sin^{2l}(t/2) cot^{2r+s-m}(t/2) = sin^{2l-p} cos^{p}, p = 2r+s-m
        p = 2 * r + s - m
        tot = tot + c * sh ** (2 * l - p) * ch ** p
    return pref * tot * np.exp(1j * m * phi)


def random_alm(rng, lmax, n):
    ls, ms = getlm(lmax)
    a = rng.standard_normal((n, len(ls))) + 1j * rng.standard_normal((n, len(ls)))
    a[:, ms == 0] = a[:, ms == 0].real
    return a


def synth_spin0(alm, lmax, th, ph):
    ls, ms = getlm(lmax)
    out = np.zeros((alm.shape[0], len(th)))
    for i, (l, m) in enumerate(zip(ls, ms)):
        y = sph_harm_y(l, m, th, ph)
        w = 1.0 if m == 0 else 2.0
        out += w * np.real(alm[:, i:i + 1] * y[None, :])
    return out


def adjoint_spin0(maps, lmax, th, ph):
    ls, ms = getlm(lmax)
    out = np.zeros((maps.shape[0], len(ls)), dtype=complex)
    for i, (l, m) in enumerate(zip(ls, ms)):
        y = sph_harm_y(l, m, th, ph)
        out[:, i] = maps @ np.conj(y)
    return out


def synth_spin(G, C, s, lmax, th, ph):
    """R + iI = -sum_{all m} (G + iC)_lm sY_lm with G_{l,-m} = (-1)^m G*_lm."""
    ls, ms = getlm(lmax)
    R = np.zeros((G.shape[0], len(th)))
    I = np.zeros((G.shape[0], len(th)))
    for i, (l, m) in enumerate(zip(ls, ms)):
        terms = [(m, G[:, i], C[:, i])]
        if m > 0:
            terms.append((-m, (-1) ** m * np.conj(G[:, i]), (-1) ** m * np.conj(C[:, i])))
        for mm, g, c in terms:
            y = goldberg_sylm(s, l, mm, th, ph)
            z = -(g + 1j * c)[:, None] * y[None, :]
            R += z.real
            I += z.imag
    # R + iI above holds only the +s combination; R, I must be real maps:
    return R, I


def spin_synthesis_matrix(s, lmax, th, ph):
    """Real matrix mapping the real parameter vector of (G, C) to stacked (R, I)."""
    ls, ms = getlm(lmax)
    cols = []
    labels = []
    for comp in range(2):
        for i, m in enumerate(ms):
            for part in range(2 if m > 0 else 1):
                e = np.zeros((1, len(ls)), dtype=complex)
                e[0, i] = 1.0 if part == 0 else 1j
                G, C = (e, 0 * e) if comp == 0 else (0 * e, e)
                R, I = synth_spin(G, C, s, lmax, th, ph)
                cols.append(np.concatenate([R[0], I[0]]))
                labels.append((comp, i, part))
    return np.array(cols).T, labels


def main():
    rng = np.random.default_rng(20250201)
    out = {}
    for nside in (2, 4, 8):
        lmax = 3 * nside - 1
        th, ph = healpix_centres(nside)
        a = random_alm(rng, lmax, 2)
        out[f"s0_n{nside}_alm"] = a
        out[f"s0_n{nside}_map"] = synth_spin0(a, lmax, th, ph)
        m = rng.standard_normal((2, len(th)))
        out[f"s0_n{nside}_inmap"] = m
        out[f"s0_n{nside}_adj"] = adjoint_spin0(m, lmax, th, ph)
    for nside in (2, 4):
        lmax = 3 * nside - 1
        th, ph = healpix_centres(nside)
        G = random_alm(rng, lmax, 1)
        C = random_alm(rng, lmax, 1)
        ls, _ = getlm(lmax)
        G[:, ls < 2] = 0
        C[:, ls < 2] = 0
        R, I = synth_spin(G, C, 2, lmax, th, ph)
        out[f"s2_n{nside}_G"], out[f"s2_n{nside}_C"] = G, C
        out[f"s2_n{nside}_R"], out[f"s2_n{nside}_I"] = R, I
        # adjoint through the explicit real synthesis matrix (its transpose)
        A, labels = spin_synthesis_matrix(2, lmax, th, ph)
        npix = len(th)
        Rin = rng.standard_normal(npix)
        Iin = rng.standard_normal(npix)
        v = A.T @ np.concatenate([Rin, Iin])
        Gadj = np.zeros(len(ls), dtype=complex)
        Cadj = np.zeros(len(ls), dtype=complex)
        ms = getlm(lmax)[1]
        for val, (comp, i, part) in zip(v, labels):
            tgt = Gadj if comp == 0 else Cadj
            # adjoint under <a,b> = sum w_m Re(a* b), w_m = 1 + [m > 0]: W^-1 A^T
            w = 1.0 if ms[i] == 0 else 2.0
            tgt[i] += (val / w) * (1.0 if part == 0 else 1j)
        out[f"s2_n{nside}_Rin"], out[f"s2_n{nside}_Iin"] = Rin, Iin
        out[f"s2_n{nside}_Gadj"], out[f"s2_n{nside}_Cadj"] = Gadj, Cadj
    np.savez_compressed(os.path.join(HERE, "sht_bruteforce.npz"), **out)

    # large-l spot values of Y_lm(theta, 0) at 60 digits
    import mpmath as mp
    mp.mp.dps = 60
    spots = [(1500, 1200, 0.6), (2048, 1500, 0.95), (2048, 1800, 1.2), (1535, 1000, 0.8),
             (1024, 1000, 1.4), (700, 3, 0.01), (2000, 1999, 1.5)]
    vals = []
    for l, m, t in spots:
        x = mp.cos(mp.mpf(t))
        p = mp.legenp(l, m, x, type=2, maxprec=200000, maxterms=10 ** 6)
        norm = mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.exp(mp.loggamma(l - m + 1) - mp.loggamma(l + m + 1)))
        vals.append(float(norm * p))
    np.savez_compressed(os.path.join(HERE, "ylm_spot.npz"), spots=np.array(spots), vals=np.array(vals))


if __name__ == "__main__":
    main()
