"""FP64 numpy oracle of the learned preconditioner's stencil network (TEST INFRASTRUCTURE ONLY).

Layer: Y[b,p,:] = act(bias + sum_{k<9} sum_c X[b, nbr[p,k], c] W[k*cin + c, :]) with
absent neighbours (nbr = -1) contributing zero; ReLU on every layer but the last,
whose output is multiplied by an optional per-pixel gate g[p].
Loss: sum_b s_b sum (Y_last - T)^2.  Gradients are written out by hand (no
autodiff) and checked against central finite differences in tests/test_nn.py.
"""
import numpy as np


def im2col(X, nbr):
    B, npix, cin = X.shape
    g = np.where(nbr[None, :, :, None] >= 0, X[:, np.maximum(nbr, 0), :], 0.)  # (B, npix, 9, cin)
    return g.reshape(B, npix, 9 * cin)


def forward(X, params, nbr, gate=None):
    """Returns the list of layer inputs (as im2col matrices) and outputs."""
    cols, outs = [], []
    h = X
    for i, (W, b) in enumerate(params):
        c = im2col(h, nbr)
        y = c @ W + b
        if i < len(params) - 1:
            y = np.maximum(y, 0.)
        elif gate is not None:
            y = y * gate[None, :, None]
        cols.append(c)
        outs.append(y)
        h = y
    return cols, outs


def loss(Y, T, s):
    return float(np.sum(s[:, None, None] * (Y - T) ** 2))


def backward(X, T, s, params, nbr, gate=None):
    """Gradients of loss(forward(X)[-1], T, s) with respect to every (W, b)."""
    cols, outs = forward(X, params, nbr, gate)
    dY = 2. * s[:, None, None] * (outs[-1] - T)
    if gate is not None:
        dY = dY * gate[None, :, None]
    grads = [None] * len(params)
    for i in range(len(params) - 1, -1, -1):
        W, b = params[i]
        if i < len(params) - 1:
            dY = dY * (outs[i] > 0)
        c = cols[i]
        B, npix, K = c.shape
        dW = c.reshape(-1, K).T @ dY.reshape(-1, dY.shape[-1])
        db = dY.reshape(-1, dY.shape[-1]).sum(0)
        grads[i] = (dW, db)
        if i > 0:
            G = dY @ W.T  # (B, npix, 9*cin)
            cin = K // 9
            G = G.reshape(B, npix, 9, cin)
            dX = np.zeros((B, npix, cin))
            for k in range(9):
                q = nbr[:, k]
                ok = q >= 0
                np.add.at(dX, (slice(None), q[ok]), G[:, ok, k, :])
            dY = dX
    return grads, outs[-1]
