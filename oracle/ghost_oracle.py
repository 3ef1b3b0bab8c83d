"""CPU oracle of the ghost-track filter (test infrastructure only).

numpy restatement, in float64, of the ten track features and of the
perceptron's forward pass and gradient as include/velo_ghost.h defines them.
The filter has no counterpart in the reference, so nothing here cites it:
the feature definitions are pinned by hand-built tracks with known answers
and the gradient by central finite differences of the loss
(tests/test_ghost_oracle.py); the GPU kernels are then compared with this
module within stated tolerances (tests/test_gpu_ghost.py).
"""
from __future__ import annotations

import numpy as np

N_FEATURES = 10
HIDDEN = 32
N_PARAMS = 1441


def unpack(params):
    p = np.asarray(params, np.float64)
    return (p[0:320].reshape(32, 10), p[320:352], p[352:1376].reshape(32, 32), p[1376:1408], p[1408:1440],
            p[1440])


def module_of(ids):
    return ((np.asarray(ids, np.uint32) >> 18) & 0xFF) >> 2


def track_features(x, y, z, ids, state):
    """Features of one track: hits in track order, state = (xb, yb, tx, ty,
    zb, chi2)."""
    xb, yb, tx, ty, zb, chi2 = (float(v) for v in state)
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    n = len(x)
    if n == 0:
        return np.array([0, 0, 0, tx, ty, chi2, 0, 0, 0, 0], np.float64)
    m = module_of(ids)
    dz = z - zb
    r = np.hypot(x - (xb + tx * dz), y - (yb + ty * dz))
    pairs = np.unique(m >> 1)
    span = int(m.max() >> 1) - int(m.min() >> 1) + 1
    phi = np.arctan2(y[:3], x[:3])
    spread = 0.0
    for i in range(len(phi)):
        for j in range(i + 1, len(phi)):
            d = abs(phi[i] - phi[j])
            if d > np.pi:
                d = 2 * np.pi - d
            spread = max(spread, d)
    return np.array([n, m.min(), m.max(), tx, ty, chi2, r.mean(), r.max(), span - len(pairs), spread], np.float64)


def consolidated_features(atomics, hit_number, track_hits, states, n_events):
    """Features of every consolidated track (reference output layout)."""
    T = int(atomics[2 * n_events])
    Ht = int(hit_number[T])
    soa = np.asarray(track_hits, np.uint8)[: 16 * Ht]
    x = soa[0: 4 * Ht].view(np.float32)
    y = soa[4 * Ht: 8 * Ht].view(np.float32)
    z = soa[8 * Ht: 12 * Ht].view(np.float32)
    ids = soa[12 * Ht: 16 * Ht].view(np.uint32)
    S = np.asarray(states, np.uint8)[: 48 * T].view(np.float32).reshape(12, T)
    out = np.zeros((T, N_FEATURES), np.float64)
    for t in range(T):
        a, b = int(hit_number[t]), int(hit_number[t + 1])
        out[t] = track_features(x[a:b], y[a:b], z[a:b], ids[a:b], (S[0, t], S[1, t], S[2, t], S[3, t], S[11, t],
                                                                     S[10, t]))
    return out


def forward(params, xn):
    """Logits and activations for normalised inputs xn (B, 10)."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    a1 = xn @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    z = h2 @ w3 + b3
    return z, (xn, a1, h1, a2, h2)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss(params, xn, y, pos_weight=1.0):
    """Mean weighted binary cross-entropy over the batch."""
    z, _ = forward(params, xn)
    w = np.where(y > 0.5, pos_weight, 1.0)
    return float(np.mean(w * (y * np.logaddexp(0, -z) + (1 - y) * np.logaddexp(0, z))))


def gradient(params, xn, y, pos_weight=1.0):
    """Analytic gradient of loss() (float64), in the parameter order."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    z, (x, a1, h1, a2, h2) = forward(params, xn)
    B = len(y)
    w = np.where(y > 0.5, pos_weight, 1.0)
    dz = w * (sigmoid(z) - y) / B
    d2 = np.outer(dz, w3) * (a2 > 0)
    d1 = (d2 @ W2) * (a1 > 0)
    g = np.concatenate([(d1.T @ x).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), h2.T @ dz, [dz.sum()]])
    return g
