"""CPU reference of the move-ordering MLP (test infrastructure only).

Only tests/ may use this module, as the checker of the GPU forward and
backward kernels (tdp_learn.hip); the product never calls it.  It restates
the same arithmetic in float64 with the same bf16 roundings as the GPU:
weights W1 and w2 rounded to bf16, the hidden pre-activations and logits
accumulated exactly, dhpre rounded to bf16 before the X^T dhpre product.
The MLP has no counterpart in the reference project (its search orders
children by triple index only, sorting_search.cu:533-535).
"""
import numpy as np

IN, HIDDEN = 128, 256


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(rows):
    """(N, 4) uint32 words -> (N, 128) 0/1 float64."""
    rows = np.asarray(rows, np.uint32)
    bits = (rows[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1
    return bits.reshape(len(rows), IN).astype(np.float64)


def _round(x, exact):
    return np.asarray(x, np.float64) if exact else bf16(x)


def forward(rows, w1, b1, w2, b2, exact=False):
    X = unpack(rows)
    hpre = X @ _round(w1, exact) + np.asarray(b1, np.float64)
    logits = np.maximum(hpre, 0.0) @ _round(w2, exact) + float(b2)
    return hpre, logits


def loss(rows, labels, group_start, w1, b1, w2, b2, exact=False):
    _, logits = forward(rows, w1, b1, w2, b2, exact)
    G = len(group_start) - 1
    out = 0.0
    for g in range(G):
        s, e = group_start[g], group_start[g + 1]
        z = logits[s:e] - logits[s:e].max()
        out += -(z[np.argmax(labels[s:e])] - np.log(np.exp(z).sum())) / G
    return out


def loss_and_gradients(rows, labels, group_start, w1, b1, w2, b2, exact=False):
    """Mean softmax cross-entropy over groups and its gradient, flat
    [w1 (k-major) | b1 | w2 | b2].  exact=True drops the bf16 roundings (for
    checks against finite differences)."""
    X = unpack(rows)
    hpre, logits = forward(rows, w1, b1, w2, b2, exact)
    G = len(group_start) - 1
    dy = np.zeros(len(logits))
    loss = 0.0
    for g in range(G):
        s, e = group_start[g], group_start[g + 1]
        z = logits[s:e] - logits[s:e].max()
        p = np.exp(z) / np.exp(z).sum()
        lab = np.asarray(labels[s:e], np.float64)
        dy[s:e] = (p - lab) / G
        loss += -np.log(p[np.argmax(lab)]) / G
    h = np.maximum(hpre, 0.0)
    mask = (hpre > 0).astype(np.float64)
    dw2 = h.T @ dy
    db2 = dy.sum()
    dh = dy[:, None] * _round(w2, exact)[None, :] * mask
    db1 = dh.sum(axis=0)
    dw1 = X.T @ _round(dh, exact)
    return loss, np.concatenate([dw1.ravel(), db1, dw2, [db2]])
