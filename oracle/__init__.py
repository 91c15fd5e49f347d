"""FP64 CPU oracle for PolySpec's hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline) -- never as the thing measured or shipped.
The product package ``polyspec_amd`` never imports it.

Contents:
  * ctypes bindings to ``liboracle.so`` (built from ``sht_oracle.c`` and
    ``kernels_oracle.c`` by ``make -C oracle``): HEALPix SHTs (ducc0's
    synthesis / adjoint_synthesis as called at polyspec/base.py:147-151,
    172-176, 193, 225, 253, 282) and the n=0 radial kernels of
    polyspec/cython/tauNL_utils.pyx and ideal_fisher.pyx;
  * numpy restatements of the small host-side formulas (base.py, weighting.py,
    tspec_template.py) used to build end-to-end oracle results.

Parity pins (tests/test_oracle_*.py): spin-0 transforms against
scipy.special.sph_harm_y brute force; spin-s transforms against the Goldberg et
al. (1967) closed form, itself checked against spin-raising (eth) derivatives of
scipy's Y_lm; adjointness; FFT against numpy.fft; kernels against brute-force
numpy einsum restatements on seeded inputs.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force=False):
    """Compile liboracle.so with gcc (make -C oracle)."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        P = ctypes.c_void_p
        i, l, d = ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.orc_sht_synthesis.argtypes = [i, i, i, i, i, P, P, i]
        L.orc_sht_adjoint.argtypes = [i, i, i, i, i, P, P, i]
        L.orc_spin_legendre.argtypes = [i, i, i, d, P]
        L.orc_ring_info.argtypes = [i, i, P, P, P, P, P]
        L.orc_apply_fl_weights.argtypes = [P, i, i, P, i, l, P, d, P]
        L.orc_gnl_loc_sum.argtypes = [P, P, P, i, l]
        L.orc_gnl_loc_sum.restype = d
        L.orc_gnl_loc_disc_sum.argtypes = [P, P, P, P, P, i, l]
        L.orc_gnl_loc_disc_sum.restype = d
        L.orc_radial_sum.argtypes = [P, i, l, P, P, i, P, P]
        L.orc_tau_sum_n0.argtypes = [P, P, i, l, i, P, P, P, P, i, i, i]
        L.orc_tau_sum_n0.restype = d
        L.orc_convolve_F_n000.argtypes = [P, l, i, P, P, P, P, P]
        L.orc_outer_product.argtypes = [P, P, i, i, l, i, P]
        L.orc_integrate_pq.argtypes = [P, i, i, P, l, P, P, P]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------------ geometry
def ring_info(nside):
    """Per-ring (z, sin theta, phi0, nphi, start) of the RING scheme (1-based ring i)."""
    L = lib()
    out = []
    for i in range(1, 4 * nside):
        z, s, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n, st = ctypes.c_int(), ctypes.c_long()
        L.orc_ring_info(nside, i, ctypes.byref(z), ctypes.byref(s), ctypes.byref(p),
                        ctypes.byref(n), ctypes.byref(st))
        out.append((z.value, s.value, p.value, n.value, st.value))
    return out


def pix_angles(nside):
    """(theta, phi) of every pixel centre in RING order."""
    npix = 12 * nside * nside
    th = np.empty(npix)
    ph = np.empty(npix)
    for z, s, p0, n, st in ring_info(nside):
        th[st:st + n] = np.arctan2(s, z)
        ph[st:st + n] = p0 + 2 * np.pi * np.arange(n) / n
    return th, ph


def nlm(lmax, mmax=None):
    mmax = lmax if mmax is None else mmax
    return (mmax + 1) * (lmax + 1) - mmax * (mmax + 1) // 2


def getlm(lmax):
    """healpy.Alm.getlm(lmax): (l, m) arrays in m-major order."""
    ls, ms = [], []
    for m in range(lmax + 1):
        ls.append(np.arange(m, lmax + 1))
        ms.append(np.full(lmax + 1 - m, m))
    return np.concatenate(ls), np.concatenate(ms)


# ---------------------------------------------------------------------- SHTs
def synthesis(alm, nside, lmax, spin=0, nthreads=0):
    """ducc0.sht.synthesis: alm (nmaps, ncomp, nlm) c128 -> maps (nmaps, ncomp, npix)."""
    alm = _c(alm, np.complex128)
    ncomp = 2 if spin else 1
    assert alm.ndim == 3 and alm.shape[1] == ncomp and alm.shape[2] == nlm(lmax)
    out = np.zeros((alm.shape[0], ncomp, 12 * nside * nside))
    rc = lib().orc_sht_synthesis(nside, lmax, lmax, spin, alm.shape[0], _p(alm), _p(out), nthreads)
    assert rc == 0
    return out


def adjoint(maps, nside, lmax, spin=0, nthreads=0):
    """ducc0.sht.adjoint_synthesis (unscaled): maps (nmaps, ncomp, npix) -> alm."""
    maps = _c(maps, np.float64)
    ncomp = 2 if spin else 1
    assert maps.ndim == 3 and maps.shape[1] == ncomp and maps.shape[2] == 12 * nside * nside
    out = np.zeros((maps.shape[0], ncomp, nlm(lmax)), dtype=np.complex128)
    rc = lib().orc_sht_adjoint(nside, lmax, lmax, spin, maps.shape[0], _p(maps), _p(out), nthreads)
    assert rc == 0
    return out


def spin_legendre(lmax, m, s, theta):
    """sY_lm(theta, 0) for l = 0..lmax from the oracle recursion."""
    out = np.zeros(lmax + 1)
    lib().orc_spin_legendre(lmax, m, s, float(theta), _p(out))
    return out


# ---- PolySpec base.py restatements on top of the oracle SHTs (FP64) ----------
def to_map(alm, nside, lmax):
    """polyspec/base.py:153-176 (ducc branch): TEB (or T) alm -> TQU (or T) maps."""
    alm = np.asarray(alm)
    if len(alm) == 3:
        out = np.zeros((3, 12 * nside * nside))
        out[0] = synthesis(alm[None, :1], nside, lmax, 0)[0, 0]
        out[1:] = synthesis(alm[None, 1:], nside, lmax, 2)[0]
        return out
    return synthesis(alm[None, :1], nside, lmax, 0)[0]


def to_lm(maps, nside, lmax):
    """polyspec/base.py:130-151 (ducc branch): x pi/(3 nside^2) = 4 pi / Npix."""
    maps = np.asarray(maps)
    sc = np.pi / (3.0 * nside ** 2)
    if len(maps) == 3:
        T = adjoint(maps[None, :1], nside, lmax, 0)[0, 0] * sc
        EB = adjoint(maps[None, 1:], nside, lmax, 2)[0] * sc
        return np.asarray([T, EB[0], EB[1]])
    return adjoint(maps[None, :1], nside, lmax, 0)[0] * sc


def to_map_vec(alm_vec, nside, lmax):
    """polyspec/base.py:225: batched spin-0 synthesis (B, nlm) -> (B, npix)."""
    return synthesis(np.asarray(alm_vec)[:, None, :], nside, lmax, 0)[:, 0]


def to_lm_vec(map_vec, nside, lmax):
    """polyspec/base.py:193: batched spin-0 adjoint x 4pi/Npix."""
    return adjoint(np.asarray(map_vec)[:, None, :], nside, lmax, 0)[:, 0] * (np.pi / (3.0 * nside ** 2))


# ------------------------------------------------------------- radial kernels
def apply_fl_weights(flX, h, ls, weight):
    flX = _c(flX, np.float64)
    h = _c(h, np.complex128)
    ls = _c(ls, np.int32)
    nr = flX.shape[2]
    out = np.zeros((nr, h.shape[1]), dtype=np.complex128)
    npol = 1 if flX.shape[1] == 1 else 2
    npol = min(npol, h.shape[0])
    lib().orc_apply_fl_weights(_p(flX), flX.shape[1], nr, _p(h), npol, h.shape[1], _p(ls), float(weight), _p(out))
    return out


def gnl_loc_sum(w, P, Q):
    w, P, Q = _c(w, np.float64), _c(P, np.float64), _c(Q, np.float64)
    return lib().orc_gnl_loc_sum(_p(w), _p(P), _p(Q), P.shape[0], P.shape[1])


def gnl_loc_disc_sum(w, P1, P2, Q1, Q2):
    a = [_c(x, np.float64) for x in (w, P1, P2, Q1, Q2)]
    return lib().orc_gnl_loc_disc_sum(*[_p(x) for x in a], a[1].shape[0], a[1].shape[1])


def radial_sum(lm_map, w, flXs, ls):
    lm_map, w, flXs, ls = _c(lm_map, np.complex128), _c(w, np.float64), _c(flXs, np.float64), _c(ls, np.int32)
    npol = flXs.shape[1]
    out = np.zeros((npol, lm_map.shape[1]), dtype=np.complex128)
    lib().orc_radial_sum(_p(lm_map), lm_map.shape[0], lm_map.shape[1], _p(w), _p(flXs), npol, _p(ls), _p(out))
    return out


def tau_sum_n0(arr1, arr2, FLLs, w, Ls, Ms, Lmin, Lmax, n1=0):
    arr1, arr2 = _c(arr1, np.complex128), _c(arr2, np.complex128)
    FLLs, w = _c(FLLs, np.float64), _c(w, np.float64)
    Ls, Ms = _c(Ls, np.int32), _c(Ms, np.int32)
    nmu, nLM, nr = arr1.shape
    return lib().orc_tau_sum_n0(_p(arr1), _p(arr2), nmu, nLM, nr, _p(FLLs), _p(w), _p(Ls), _p(Ms),
                                int(Lmin), int(Lmax), int(n1))


def convolve_F_n000(PQ, FLL, w, Ls, L_indices, nout):
    PQ, FLL, w = _c(PQ, np.complex128), _c(FLL, np.float64), _c(w, np.float64)
    Ls, Li = _c(Ls, np.int32), _c(L_indices, np.int32)
    nLM, nr = PQ.shape
    out = np.zeros((nout, nr), dtype=np.complex128)
    lib().orc_convolve_F_n000(_p(PQ), nLM, nr, _p(FLL), _p(w), _p(Ls), _p(Li), _p(out))
    return out


def outer_product(Qa, Qb, sym=False):
    Qa, Qb = _c(Qa, np.complex128), _c(Qb, np.complex128)
    fish = np.zeros((Qa.shape[1], Qb.shape[1]))
    lib().orc_outer_product(_p(Qa), _p(Qb), Qa.shape[1], Qb.shape[1], Qa.shape[2], int(sym), _p(fish))
    return fish


def integrate_pq(flXs, P, ls, w, out):
    flXs, P, ls, w = _c(flXs, np.float64), _c(P, np.complex128), _c(ls, np.int32), _c(w, np.float64)
    assert out.dtype == np.complex128 and out.flags.c_contiguous
    lib().orc_integrate_pq(_p(flXs), flXs.shape[1], P.shape[1], _p(P), P.shape[0], _p(ls), _p(w), _p(out))
    return out
