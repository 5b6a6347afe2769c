"""CPU oracle for the learned parts (TEST INFRASTRUCTURE ONLY).

float64 numpy restatements of
  * the per-vertex features (pumatch_amd/csrc/learn.hip, k_features),
  * the 16 -> 64 -> 64 -> 2 ReLU MLP: forward, the weighted squared-error
    loss and its gradient (k_mlp_grad), and inference with BF16-rounded
    weights and inputs (k_mlp_infer).
The reference has no learned component, so nothing here follows a
reference file; the oracle is pinned instead by central finite differences
of its own loss (tests/test_learn_oracle.py), by a hand-computed one-sample
case, and by closed forms of the features on complete and star graphs.
Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

F, H, O = 16, 64, 2
OFF = {}
_o = 0
for _name, _size in (("W1", F * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * O), ("b3", O)):
    OFF[_name] = (_o, _size)
    _o += _size
NPARAMS = _o  # 5378


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    get = lambda k: w[OFF[k][0]:OFF[k][0] + OFF[k][1]]
    return (get("W1").reshape(F, H), get("b1"), get("W2").reshape(H, H), get("b2"),
            get("W3").reshape(H, O), get("b3"))


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(x) for x in (W1, b1, W2, b2, W3, b3)])


def forward(w, X):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = np.asarray(X, dtype=np.float64).reshape(-1, F)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3 + b3
    return y, (X, z1, h1, z2, h2)


def loss_grad(w, X, Y, Wt=None):
    """loss = sum_{s,o} Wt (yhat - y)^2 / sum Wt and d loss / d params."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    y, (Xd, z1, h1, z2, h2) = forward(w, X)
    Y = np.asarray(Y, dtype=np.float64).reshape(-1, O)
    Wt = np.ones_like(Y) if Wt is None else np.asarray(Wt, dtype=np.float64).reshape(-1, O)
    scale = 1.0 / Wt.sum()
    r = y - Y
    loss = scale * float((Wt * r * r).sum())
    d3 = 2.0 * scale * Wt * r
    gW3 = h2.T @ d3
    gb3 = d3.sum(0)
    d2 = (d3 @ W3.T) * (z2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = Xd.T @ d1
    gb1 = d1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def bf16(x):
    """Round float values to bfloat16 (nearest, ties to even), as floats."""
    a = np.asarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(w, X):
    """Inference as the MFMA kernel computes it: weights, inputs and the
    hidden activations rounded to bf16, products summed in float."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = bf16(np.asarray(X, dtype=np.float32).reshape(-1, F))
    h1 = bf16(np.maximum(X @ bf16(W1) + b1, 0))
    h2 = bf16(np.maximum(h1 @ bf16(W2) + b2, 0))
    return h2 @ bf16(W3) + b3


def features(rowptr, colidx, labels=None):
    """Per-vertex features, the same definitions as the GPU kernel."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    colidx = np.asarray(colidx, dtype=np.int64)
    n = rowptr.size - 1
    deg = np.diff(rowptr)
    maxnbr = np.zeros(n, np.int64)
    out = np.zeros((n, F), np.float64)
    for v in range(n):
        nv = colidx[rowptr[v]:rowptr[v + 1]]
        d = nv.size
        if d:
            maxnbr[v] = deg[nv].max()
        two = float(deg[nv].sum())
        same = int((labels[nv] == labels[v]).sum()) if labels is not None else 0
        tri = 0.0
        if d >= 2:
            ns = min(d, 4)
            hit = 0.0
            for s in range(ns):
                w = nv[s * d // ns]
                nw = colidx[rowptr[w]:rowptr[w + 1]]
                a, b = (nv, nw) if d <= nw.size else (nw, nv)
                na = a.size
                take = min(na, 64)
                bs = set(b.tolist())
                c = sum(1 for i in range(take) if int(a[i * na // take]) in bs)
                hit += c * na / take
            tri = hit / ns * d / 2.0
        out[v, 0] = np.log2(1.0 + d)
        out[v, 1] = np.log2(1.0 + two)
        out[v, 2] = np.log2(1.0 + maxnbr[v])
        out[v, 3] = np.log2(1.0 + tri)
        out[v, 4] = (v & ~4095) / n
        out[v, 5] = np.log2(1.0 + v) / np.log2(1.0 + n)
        if labels is not None:
            out[v, 6 + (int(labels[v]) & 7)] = 1.0
            out[v, 14] = np.log2(1.0 + same)
        out[v, 15] = 1.0
    return out
