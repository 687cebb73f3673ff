"""CPU oracle of the node-scoring network and of the node features.

TEST INFRASTRUCTURE ONLY: the checker the GPU network (zg_mlp_*) and feature
kernel (zg_node_features) are compared against; the product never imports
it.  Neither has a counterpart in the reference: the network is the learned
pivot scorer beside Brandes, and its features are computed from the graph
(degrees, a 2-hop degree sum, PageRank, and core numbers as
priv/zog/src/connectivity.zig:5-101 defines them).

The network is 8 -> 128 -> 128 -> 1 with ReLU hidden layers; parameters are
one float32 vector W1[8][128], b1[128], W2[128][128], b2[128], w3[128], b3.
`forward`/`gradient` with bf16=True reproduce the GPU's roundings: BF16
operands of the three matrix products per pass (x, W1, relu(z1), W2 in the
forward pass; dz2 and W2 for dh1; relu(z1), dz2, x and dz1 for the weight
gradients) and exact (here FP64) accumulation; the other terms are kept in
full precision.  With bf16=False it is the plain FP64 network, which
`tests/test_mlp_oracle.py` pins against central finite differences.
"""
from __future__ import annotations

import numpy as np

NFEAT = 8
H = 128
P = NFEAT * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NFEAT * H].reshape(NFEAT, H); o += NFEAT * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]]).astype(np.float64)


def bf16(a):
    """Round to BF16 (nearest even) and back, as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def forward(p, x, bf16_ops=True):
    """y[n] and the intermediates (z1, h1, z2) of the network on x[n][8]."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    xr = _r(x, bf16_ops)
    z1 = xr @ _r(W1, bf16_ops) + b1
    h1 = _r(np.maximum(z1, 0.0), bf16_ops)
    z2 = h1 @ _r(W2, bf16_ops) + b2
    y = np.maximum(z2, 0.0) @ w3 + b3
    return y, (xr, z1, h1, z2)


def gradient(p, x, t, bf16_ops=True):
    """(loss, grad[P]) of the mean squared error (1/n) sum (y - t)^2."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = len(x)
    y, (xr, z1, h1, z2) = forward(p, x, bf16_ops)
    err = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(err ** 2))
    dy = (2.0 / n) * err
    a2 = np.maximum(z2, 0.0)
    gw3 = a2.T @ dy
    gb3 = dy.sum()
    dz2 = dy[:, None] * w3[None, :] * (z2 > 0)
    gb2 = dz2.sum(0)
    dz2r = _r(dz2, bf16_ops)
    dh1 = dz2r @ _r(W2, bf16_ops).T
    dz1 = dh1 * (h1 != 0)
    gb1 = dz1.sum(0)
    gW2 = h1.T @ dz2r
    gW1 = xr.T @ _r(dz1, bf16_ops)
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3)


def features(n_cap, src, dst, live, pagerank, core):
    """Standardised node features [n_cap][8] as zg_node_features defines them
    (pagerank: the graph's PageRank vector; core: its core numbers)."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    live = np.ones(n_cap, bool) if live is None else np.asarray(live, bool)
    keep = live[src] & live[dst]
    s, d = src[keep], dst[keep]
    od = np.bincount(s, minlength=n_cap).astype(np.float64)
    idg = np.bincount(d, minlength=n_cap).astype(np.float64)
    two = np.bincount(s, weights=od[d], minlength=n_cap)
    n_live = int(live.sum())
    f = np.zeros((n_cap, NFEAT))
    f[:, 0] = np.log1p(od + idg)
    f[:, 1] = np.log1p(idg)
    f[:, 2] = np.log1p(od)
    f[:, 3] = np.log1p(two)
    f[:, 4] = np.log(np.maximum(np.asarray(pagerank, np.float64) * n_live, 1e-30))
    f[:, 5] = np.log1p(np.asarray(core, np.float64))
    f[:, 6] = 1.0
    for k in range(6):
        col = f[live, k]
        mean = col.mean() if n_live else 0.0
        var = max((col ** 2).mean() - mean ** 2, 0.0) if n_live else 0.0
        sd = np.sqrt(var) if var > 1e-12 else 1.0
        f[:, k] = (f[:, k] - mean) / sd
    f[~live] = 0.0
    return f
