"""Host (FP64) precompute of the radial profiles used by the local templates.

Restates polyspec/cython/k_integrals.pyx for the n = 0 (local) case:
  * p_integral (k_integrals.pyx:262-309):  p_l^X(r) = (-1)^l (2/pi) int dk k^2 P_zeta(k) T^X_l(k) j_l(k r)
  * q_integral (k_integrals.pyx:312-352):  q_l^X(r) = (-1)^l (2/pi) int dk k^2 T^X_l(k) j_l(k r)
    (trapezoid rule on the input k grid, exactly as the Cython sums
    sum_k (k_i - k_{i-1}) (f_i + f_{i-1}) with the 1/2 folded into kprod);
  * compute_F_integral (k_integrals.pyx:453-494) for n = 0: the closed form
    F_L(r, r') = pref 2^(ns-3) Gamma(L+(ns-1)/2) / (Gamma(1-(ns-2)/2) Gamma(L+3/2))
                 2F1(L+(ns-1)/2, (ns-2)/2; L+3/2; r_<^2/r_>^2) (r_</r_>)^L r_>^(1-ns),
    with 2F1 evaluated by hyp2f1_F (GSL's gsl_sf_hyperg_2F1 in the reference).
Spherical Bessel functions are evaluated directly at k r with
scipy.special.spherical_jn (the reference tabulates j_l on an x grid with GSL's
Steed algorithm and interpolates linearly, k_integrals.pyx:28-110; direct
evaluation removes that interpolation error).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
from scipy.special import gammaln, gammasgn, hyp2f1, spherical_jn


def Pzeta(k, As=2.1e-9, ns=0.96, k_pivot=0.05):
    """Primordial spectrum 2 pi^2/k^3 A_s (k/k_p)^(ns-1) (tspec_template.py:71)."""
    return 2. * np.pi ** 2 / k ** 3 * As * (k / k_pivot) ** (ns - 1)


def jl_kr(ls, k_arr, r_arr, nthreads=None):
    """j_l(k r) as (len(ls), len(r), len(k)); the values of l are spread over a thread pool
    (scipy's ufunc releases the GIL on whole arrays)."""
    ls = np.asarray(ls)
    x = np.asarray(r_arr)[:, None] * np.asarray(k_arr)[None, :]
    out = np.empty((len(ls), len(r_arr), len(k_arr)))

    def one(i):
        out[i] = spherical_jn(int(ls[i]), x)
    n = nthreads or min(16, os.cpu_count() or 1)
    if n > 1 and len(ls) > 1:
        with ThreadPoolExecutor(n) as ex:
            list(ex.map(one, range(len(ls))))
    else:
        for i in range(len(ls)):
            one(i)
    return out


def _trapz_sum(k_arr, integrand):
    """sum_i (k_i - k_{i-1}) (f_i + f_{i-1}) / 2 along the last axis."""
    dk = np.diff(k_arr)
    return np.sum(dk * (integrand[..., 1:] + integrand[..., :-1]), axis=-1) / 2.


def q_integral(k_arr, Tl_arr, r_arr, lmin, lmax, nslot, jl=None):
    """qlXs (lmax+1, nslot, N_r): zero outside lmin <= l <= lmax and X >= len(Tl_arr).
    jl: j_l(k r) for l = lmin..lmax from jl_kr, if already computed."""
    k_arr = np.asarray(k_arr, dtype=np.float64)
    ls = np.arange(lmin, lmax + 1)
    if jl is None:
        jl = jl_kr(ls, k_arr, r_arr)  # (nl, nr, nk)
    out = np.zeros((lmax + 1, nslot, len(r_arr)))
    kp = 2. / np.pi * k_arr ** 2
    sgn = (-1.) ** ls
    for X in range(len(Tl_arr)):
        f = kp[None, None, :] * np.asarray(Tl_arr[X])[lmin:lmax + 1, None, :] * jl
        out[lmin:lmax + 1, X, :] = sgn[:, None] * _trapz_sum(k_arr, f)
    return out


def p_integral_n0(k_arr, Pz, Tl_arr, r_arr, lmin, lmax, lmax_arr, nslot, jl=None):
    """plLXs[nmax=0] (lmax_arr+1, nslot, N_r): zero outside lmin <= l <= lmax."""
    k_arr = np.asarray(k_arr, dtype=np.float64)
    ls = np.arange(lmin, lmax + 1)
    if jl is None:
        jl = jl_kr(ls, k_arr, r_arr)
    out = np.zeros((lmax_arr + 1, nslot, len(r_arr)))
    kp = 2. / np.pi * k_arr ** 2 * np.asarray(Pz)
    sgn = (-1.) ** ls
    for X in range(len(Tl_arr)):
        f = kp[None, None, :] * np.asarray(Tl_arr[X])[lmin:lmax + 1, None, :] * jl
        out[lmin:lmax + 1, X, :] = sgn[:, None] * _trapz_sum(k_arr, f)
    return out


def _hyp2f1_near_one(a, b, c, w):
    """2F1(a, b; c; 1 - w) for small a w by the transformation to 1 - z (Abramowitz & Stegun
    15.3.6, s = c - a - b not an integer):
      Gamma(c) Gamma(s) / (Gamma(c-a) Gamma(c-b)) 2F1(a, b; 1-s; w)
      + w^s Gamma(c) Gamma(-s) / (Gamma(a) Gamma(b)) 2F1(c-a, c-b; 1+s; w),
    both series summed directly (their terms stay below e^(a w)); Gamma ratios in log form.
    w: array >= 0.  w = 0 gives the Gauss sum Gamma(c) Gamma(s) / (Gamma(c-a) Gamma(c-b))."""
    s = c - a - b
    w = np.asarray(w, dtype=np.float64)

    def series(p, q, r):
        tot = np.ones_like(w)
        t = np.ones_like(w)
        for n in range(400):
            t = t * (p + n) * (q + n) / ((r + n) * (n + 1.)) * w
            tot = tot + t
            if np.all(np.abs(t) <= 1e-17 * np.abs(tot)):
                break
        return tot
    g1 = gammasgn(s) * gammasgn(c - a) * gammasgn(c - b) * np.exp(gammaln(c) + gammaln(s) - gammaln(c - a) -
                                                                 gammaln(c - b))
    out = g1 * series(a, b, 1. - s)
    pos = w > 0
    if np.any(pos):
        g2 = gammasgn(-s) * gammasgn(a) * gammasgn(b) * np.exp(gammaln(c) + gammaln(-s) - gammaln(a) - gammaln(b))
        out[pos] += w[pos] ** s * g2 * series(c - a, c - b, 1. + s)[pos]
    return out


def hyp2f1_F(a, b, c, z):
    """2F1(a, b; c; z) for 0 <= z <= 1 and c - a - b > 0 non-integer, as the F_L closed form needs
    it up to large L (a ~ c ~ L): scipy's hyp2f1 overflows its Gamma ratios there once z is within
    about 2e-3 of 1 (inf / nan for L >~ 170, even 0 at z = 1, L = 170).  Within 4e-3 of 1 (where
    a (1 - z) < 5, so that the series terms stay small) the transformation to 1 - z is used, scipy
    elsewhere, and mpmath for anything left non-finite."""
    z = np.asarray(z, dtype=np.float64)
    w = 1. - z
    near = (w < 4e-3) & (a * w < 5.)
    out = np.empty_like(z)
    if np.any(near):
        out[near] = _hyp2f1_near_one(a, b, c, w[near])
    if np.any(~near):
        out[~near] = hyp2f1(a, b, c, z[~near])
    bad = ~np.isfinite(out)
    if np.any(bad):
        import mpmath
        with mpmath.workdps(30):
            out[bad] = [float(mpmath.hyp2f1(a, b, c, float(x))) for x in z[bad]]
    return out


def F_integral_n0(r_arr, Lmin, Lmax, ns, As, k_pivot):
    """FLLs[nmax_F = 0] (Lmax+1, N_r, N_r), zero for L < max(Lmin, 1)."""
    r = np.asarray(r_arr, dtype=np.float64)
    nr = len(r)
    out = np.zeros((Lmax + 1, nr, nr))
    pref = 2. ** (ns - 3) * 2. * np.pi ** 2 * As / k_pivot ** (ns - 1)
    i2 = (ns - 2.) / 2.
    rl = np.minimum(r[:, None], r[None, :])
    rg = np.maximum(r[:, None], r[None, :])
    ratio = rl / rg
    for L in range(max(Lmin, 1), Lmax + 1):
        i1 = L + (ns - 1.) / 2.
        i3 = L + 1.5
        pre = pref * np.exp(gammaln(i1) - gammaln(1. - i2) - gammaln(i3))
        out[L] = pre * hyp2f1_F(i1, i2, i3, ratio ** 2) * ratio ** L * rg ** (1. - ns)
    return out
