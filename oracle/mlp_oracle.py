"""FP64 CPU oracle of the proposal network -- TEST INFRASTRUCTURE ONLY.

Forward and backward of d -> H1 -> H2 -> C with GELU (erf form) and mean softmax
cross-entropy against soft targets (uniform over each row's labels; one label = one-hot),
in float64 numpy, and the list proposal taken from the logits (propose_lists).  With round16=True the inputs, weights and the two
hidden activations are rounded to FP16 first, as the GPU kernels store them, so the
remaining difference to the GPU is FP16 rounding of the back-propagated terms and FP32
accumulation.  Parameter layout: W1[H1][D], b1, W2[H2][H1], b2, W3[C][H2], b3
(concatenated), the library's order.  No reference counterpart (the reference has no
learned component).
"""
from __future__ import annotations

import math

import numpy as np

_erf = np.vectorize(math.erf)


def split(params, D, H1, H2, C):
    sizes = [H1 * D, H1, H2 * H1, H2, C * H2, C]
    out, o = [], 0
    for s in sizes:
        out.append(np.asarray(params[o:o + s], np.float64))
        o += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(H1, D), b1, W2.reshape(H2, H1), b2, W3.reshape(C, H2), b3


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def _r16(a, on):
    return np.asarray(a, np.float16).astype(np.float64) if on else np.asarray(a, np.float64)


def forward(params, x, dims, round16=True):
    D, H1, H2, C = dims
    W1, b1, W2, b2, W3, b3 = split(params, *dims)
    x = _r16(x, round16)
    W1, W2, W3 = _r16(W1, round16), _r16(W2, round16), _r16(W3, round16)
    z1 = x @ W1.T + b1
    h1 = _r16(gelu(z1), round16)
    z2 = h1 @ W2.T + b2
    h2 = _r16(gelu(z2), round16)
    z3 = h2 @ W3.T + b3
    return {"x": x, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "z3": z3, "W": (W1, W2, W3)}


def targets(labels, B, C):
    """Soft targets: row i uniform over its labels below C (labels [B] or [B, L]; a label
    >= C is no label, a row without any has an all-zero target)."""
    lab = np.asarray(labels, np.int64).reshape(B, -1)
    t = np.zeros((B, C))
    for i in range(B):
        v = [int(a) for a in lab[i] if 0 <= a < C]
        for a in v:
            t[i, a] += 1.0 / len(v)
    return t


def loss_grad(params, x, labels, dims, round16=True):
    """Mean soft-target cross-entropy sum_j -t_ij log p_ij over the B rows and its gradient
    (labels [B] is the one-hot case)."""
    D, H1, H2, C = dims
    f = forward(params, x, dims, round16)
    B = x.shape[0]
    z3 = f["z3"]
    mx = z3.max(1, keepdims=True)
    e = np.exp(z3 - mx)
    se = e.sum(1, keepdims=True)
    p = e / se
    t = targets(labels, B, C)
    has = t.sum(1) > 0
    logp = z3 - mx - np.log(se)
    loss = float(-(t * logp).sum() / B)
    dz3 = np.where(has[:, None], p - t, 0.0)
    W1, W2, W3 = f["W"]
    gW3 = dz3.T @ f["h2"] / B
    gb3 = dz3.sum(0) / B
    dh2 = dz3 @ W3
    dz2 = dh2 * gelu_grad(f["z2"])
    gW2 = dz2.T @ f["h1"] / B
    gb2 = dz2.sum(0) / B
    dh1 = dz2 @ W2
    dz1 = dh1 * gelu_grad(f["z1"])
    gW1 = dz1.T @ f["x"] / B
    gb1 = dz1.sum(0) / B
    grads = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grads


def finite_difference(params, x, labels, dims, idx, h=1e-6):
    """Central differences of the FP64 (unrounded) loss for the parameter indices idx."""
    out = []
    for i in idx:
        pp = np.array(params, np.float64)
        pm = pp.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = loss_grad(pp, x, labels, dims, round16=False)
        lm, _ = loss_grad(pm, x, labels, dims, round16=False)
        out.append((lp - lm) / (2 * h))
    return np.array(out)


def propose_lists(logits, P, mass):
    """The P largest logits of each row (descending, lowest index among equal values) and
    how many of them it takes for their softmax probabilities to sum to at least `mass`
    (0 if all P fall short), in float64."""
    z = np.asarray(logits, np.float64)
    B, C = z.shape
    order = np.lexsort((np.tile(np.arange(C), (B, 1)), -z), axis=1)[:, :P]
    p = np.exp(z - z.max(1, keepdims=True))
    p /= p.sum(1, keepdims=True)
    cum = np.cumsum(np.take_along_axis(p, order, 1), 1)
    reach = cum >= mass
    counts = np.where(reach.any(1), reach.argmax(1) + 1, 0)
    return order.astype(np.uint32), counts.astype(np.uint32), cum
