"""ORACLE (test infrastructure only): CPU FP64 restatement of the learned row
binning -- the 16 row features, the 16 -> 32 -> 6 MLP's forward pass, mean
cross-entropy and its gradient, and the bin decision rule.

This part has no reference counterpart (the reference has no predictor); the
oracle states the library's own definitions so the GPU kernels can be
checked.  Pinning: the gradient against central finite differences of the
loss (tests/test_mlp_oracle.py), the features against hand-worked values of a
small product, the bins against their definition.

Parameter layout (742): W1[32][16] | b1[32] | W2[6][32] | b2[6].
"""
from __future__ import annotations

import numpy as np

NIN, NHID, NOUT = 16, 32, 6
NPARAM = NHID * NIN + NHID + NOUT * NHID + NOUT
BIN_MAX = (32, 128, 512, 2048, 4096)


def bin_of(nnz):
    nnz = np.asarray(nnz)
    out = np.full(nnz.shape, len(BIN_MAX), np.int32)
    for b in reversed(range(len(BIN_MAX))):
        out[nnz <= BIN_MAX[b]] = b
    return out


def unpack(P):
    P = np.asarray(P, np.float64)
    o = 0
    W1 = P[o:o + NHID * NIN].reshape(NHID, NIN)
    o += NHID * NIN
    b1 = P[o:o + NHID]
    o += NHID
    W2 = P[o:o + NOUT * NHID].reshape(NOUT, NHID)
    o += NOUT * NHID
    b2 = P[o:o + NOUT]
    return W1, b1, W2, b2


def forward(P, X):
    """Logits Z[n][6] (FP64)."""
    W1, b1, W2, b2 = unpack(P)
    X = np.asarray(X, np.float64)
    pre = X @ W1.T + b1
    h = np.maximum(pre, 0.0)
    return h @ W2.T + b2


def loss_and_grad(P, X, Y):
    """Mean cross-entropy over the rows and its gradient (FP64, analytic
    backpropagation)."""
    W1, b1, W2, b2 = unpack(P)
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y)
    n = len(X)
    pre = X @ W1.T + b1
    h = np.maximum(pre, 0.0)
    z = h @ W2.T + b2
    zmax = z.max(axis=1, keepdims=True)
    e = np.exp(z - zmax)
    s = e.sum(axis=1, keepdims=True)
    p = e / s
    lse = (zmax + np.log(s))[:, 0]
    loss = float(np.mean(lse - z[np.arange(n), Y]))
    dz = p.copy()
    dz[np.arange(n), Y] -= 1.0
    dz /= n
    gW2 = dz.T @ h
    gb2 = dz.sum(axis=0)
    dh = dz @ W2
    dpre = dh * (pre > 0)
    gW1 = dpre.T @ X
    gb1 = dpre.sum(axis=0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2])


def predict_bins(P, X, tau):
    """Smallest bin whose cumulative softmax probability reaches tau."""
    z = forward(P, X)
    z = z - z.max(axis=1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=1, keepdims=True)
    cum = np.cumsum(p, axis=1)
    out = np.argmax(cum >= tau, axis=1)
    out[~(cum >= tau).any(axis=1)] = NOUT - 1
    return out


def _lg(x):
    return np.log2(1.0 + x)


def row_features(Lp, Li, Ap, Ai, n):
    """Features of every output row of C = L A (row i walks L's row i; each
    entry k names a row of A).  Same definitions as the device kernel."""
    Lp, Li, Ap, Ai = (np.asarray(a, np.int64) for a in (Lp, Li, Ap, Ai))
    m = len(Lp) - 1
    F = np.zeros((m, NIN))
    for r in range(m):
        ks = Li[Lp[r]:Lp[r + 1]]
        deg = len(ks)
        lens = (Ap[ks + 1] - Ap[ks]).astype(np.float64)
        ne = lens > 0
        terms = lens.sum()
        maxlen = lens.max() if deg else 0.0
        minlen = lens[ne].min() if ne.any() else 0.0
        sumsq = (lens ** 2).sum()
        if ne.any():
            firsts = Ai[Ap[ks[ne]]]
            lasts = Ai[Ap[ks[ne] + 1] - 1]
            span = float(lasts.max() - firsts.min() + 1)
            sspan = float((lasts - firsts + 1).sum())
        else:
            span = sspan = 0.0
        kspan = float(ks.max() - ks.min() + 1) if deg else 0.0
        mean = terms / deg if deg else 0.0
        f = F[r]
        f[0] = _lg(terms)
        f[1] = _lg(deg)
        f[2] = _lg(maxlen)
        f[3] = _lg(minlen)
        f[4] = _lg(mean)
        f[5] = _lg(span)
        f[6] = _lg(span * (1 - np.exp(-terms / span)) if span > 0 else 0.0)
        f[7] = _lg(sspan)
        f[8] = _lg(min(terms, sspan))
        f[9] = _lg(min(terms, span))
        f[10] = _lg(kspan)
        f[11] = deg / kspan if kspan > 0 else 0.0
        f[12] = _lg(np.sqrt(sumsq))
        f[13] = _lg(sspan / deg if deg else 0.0)
        f[14] = _lg(maxlen) - _lg(mean)
        f[15] = _lg(terms / span if span > 0 else 0.0)
    return F
