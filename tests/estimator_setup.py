"""Shared synthetic inputs for the estimator tests (the reduced-size analogue of
the Nside=64 / lmax=128 / 10-node workload): CMB-like spectra C_l = 2 pi 1e3 /
(l(l+1)) with a Gaussian beam and white noise, transfer T_l(k) = j_l(k * 14000 Mpc)
on log-spaced k in [1e-4, 0.2] / Mpc, Gauss-Legendre radial nodes on
[13400, 14600] Mpc with weights r^2 w_GL."""
import numpy as np
from scipy.special import roots_legendre, spherical_jn


def spectra(nside, pol, fwhm_arcmin=30., noise_uK_arcmin=30.):
    L = 3 * nside - 1
    l = np.arange(L + 1)
    cl = np.zeros(L + 1)
    cl[2:] = 2 * np.pi * 1e3 / (l[2:] * (l[2:] + 1.))
    sig = np.radians(fwhm_arcmin / 60.) / np.sqrt(8 * np.log(2))
    b = np.exp(-l * (l + 1) * sig ** 2 / 2)
    nl = (noise_uK_arcmin * np.pi / 10800.) ** 2 * np.ones(L + 1)
    if not pol:
        raw = {'TT': cl}
        tot = {'TT': cl * b ** 2 + nl}
        return raw, tot, [b]
    raw = {'TT': cl, 'EE': 0.05 * cl, 'TE': 0.3 * np.sqrt(0.05) * cl, 'BB': 0. * cl}
    tot = {'TT': cl * b ** 2 + nl, 'EE': 0.05 * cl * b ** 2 + 2 * nl, 'TE': 0.3 * np.sqrt(0.05) * cl * b ** 2,
           'BB': 2 * nl}
    return raw, tot, [b, b]


def transfer(lmax, pol, nk=60):
    k = np.geomspace(1e-4, 0.2, nk)
    Tl = np.zeros((lmax + 1, nk))
    for l in range(2, lmax + 1):
        Tl[l] = spherical_jn(l, k * 14000.)
    return k, (np.asarray([Tl, Tl]) if pol else Tl[None])


def radial_nodes(n):
    x, w = roots_legendre(n)
    a, b = 13400., 14600.
    r = 0.5 * (b - a) * x + 0.5 * (b + a)
    return r, r ** 2 * w * 0.5 * (b - a)


def random_alm(rng, lmax, shape):
    from oracle import getlm
    ls, ms = getlm(lmax)
    a = rng.standard_normal(shape + (len(ls),)) + 1j * rng.standard_normal(shape + (len(ls),))
    a[..., ms == 0] = a[..., ms == 0].real * np.sqrt(2)
    return a / np.sqrt(2)


def grf_alm(rng, Cl, lmax, pol):
    """Seeded Gaussian alm with the given total spectra (T or T,E,B)."""
    from oracle import getlm
    ls, ms = getlm(lmax)
    nf = 3 if pol else 1
    names = [['TT', 'TE', 'TB'], ['TE', 'EE', 'EB'], ['TB', 'EB', 'BB']]
    C = np.zeros((lmax + 1, nf, nf))
    for a in range(nf):
        for b in range(nf):
            k = names[a][b]
            if k in Cl:
                C[:, a, b] = Cl[k][:lmax + 1]
    Lc = np.zeros_like(C)
    for l in range(lmax + 1):
        w, v = np.linalg.eigh(C[l])
        Lc[l] = v * np.sqrt(np.clip(w, 0, None))[None, :]
    z = random_alm(rng, lmax, (nf,))
    return np.einsum('lab,bl->al', Lc[ls], z)
