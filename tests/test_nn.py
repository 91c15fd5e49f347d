"""Learned-preconditioner network on the CPU: the stencil tables against brute-force
geometry, and the FP64 oracle's hand-written gradients against central finite
differences (so a dropped ReLU mask, a transposed weight or a wrong reverse-table
entry fails here before the GPU kernels are compared with the oracle)."""
import numpy as np
import pytest

import oracle
from oracle import nn as onn
from polyspec_amd.healpix import stencil9
from polyspec_amd.learned import lm_select


@pytest.mark.parametrize("nside", [2, 4, 8])
def test_stencil_geometry(nside):
    nbr, rev = stencil9(nside)
    npix = 12 * nside ** 2
    assert nbr.shape == (npix, 9) and (nbr[:, 0] == np.arange(npix)).all()
    th, ph = oracle.pix_angles(nside)
    v = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    ok = nbr >= 0
    # seven neighbours at the 8 three-face vertices (3 pixels each)
    assert (~ok).sum() == 24
    d = np.arccos(np.clip(np.einsum('pi,pki->pk', v, v[np.maximum(nbr, 0)]), -1, 1))
    assert d[ok].max() < 2.5 * np.sqrt(4 * np.pi / npix)
    # the four nearest pixels (brute force) are always in the stencil, with no repeats
    D = np.arccos(np.clip(v @ v.T, -1, 1))
    np.fill_diagonal(D, 9.)
    near = np.argsort(D, 1)[:, :4]
    for p in range(npix):
        s = nbr[p][nbr[p] >= 0]
        assert len(set(s)) == len(s)
        assert set(near[p]) <= set(s)


def test_reverse_table():
    nbr, rev = stencil9(4)
    slots = [[] for _ in range(nbr.shape[0])]
    for p in range(nbr.shape[0]):
        for k in range(9):
            if nbr[p, k] >= 0:
                slots[nbr[p, k]].append(p * 9 + k)
    for q in range(nbr.shape[0]):
        got = sorted(rev[q][rev[q] >= 0].tolist())
        assert got == sorted(slots[q])


def test_lm_select():
    from oracle import getlm
    L, Lo = 11, 5
    l, m = getlm(L)
    sel = lm_select(L, Lo)
    lo_l, lo_m = getlm(Lo)
    assert (l[sel] == lo_l).all() and (m[sel] == lo_m).all()


def _net(rng, ch, scale=0.3):
    return [(rng.standard_normal((9 * ch[i], ch[i + 1])) * scale, rng.standard_normal(ch[i + 1]) * 0.1)
            for i in range(len(ch) - 1)]


@pytest.mark.parametrize("gated", [False, True])
def test_oracle_gradients_finite_difference(gated):
    rng = np.random.default_rng(1)
    nbr, _ = stencil9(2)
    ch = [4, 5, 5, 2]
    params = _net(rng, ch)
    X = rng.standard_normal((2, nbr.shape[0], ch[0]))
    T = rng.standard_normal((2, nbr.shape[0], ch[-1]))
    s = np.array([0.7, 1.3])
    gate = rng.uniform(0, 1, nbr.shape[0]) if gated else None
    grads, _ = onn.backward(X, T, s, params, nbr, gate)
    h = 1e-6
    for i, (W, b) in enumerate(params):
        for arr, g in ((W, grads[i][0]), (b, grads[i][1])):
            flat = arr.reshape(-1)
            for j in rng.choice(flat.size, size=min(12, flat.size), replace=False):
                old = flat[j]
                flat[j] = old + h
                lp = onn.loss(onn.forward(X, params, nbr, gate)[1][-1], T, s)
                flat[j] = old - h
                lm = onn.loss(onn.forward(X, params, nbr, gate)[1][-1], T, s)
                flat[j] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g.reshape(-1)[j]) < 1e-6 * max(1., abs(fd)), (i, j, fd, g.reshape(-1)[j])


def test_oracle_im2col_is_stencil_sum():
    """A layer with one-hot weights returns the neighbour values themselves."""
    rng = np.random.default_rng(2)
    nbr, _ = stencil9(2)
    X = rng.standard_normal((1, nbr.shape[0], 1))
    for k in range(9):
        W = np.zeros((9, 1))
        W[k, 0] = 1.
        Y = onn.forward(X, [(W, np.zeros(1))], nbr)[1][0][0, :, 0]
        want = np.where(nbr[:, k] >= 0, X[0, np.maximum(nbr[:, k], 0), 0], 0.)
        assert np.array_equal(Y, want)
