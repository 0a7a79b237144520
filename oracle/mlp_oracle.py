"""CPU oracle of the learned pair-cost model (test infrastructure only).

Only tests, the package's smoke check and bench.py's CPU-baseline leg use this
module, as the checker of the device kernels in
biwfa_gpu_pii2025_amd/csrc/costmodel.hip; nothing in the product path imports it.

The cost model has no counterpart in the reference (its scheduler is a plain
loop over 100-pair launches, src/biWFA.cu:935-1004).  What it computes is fixed
by its definition:

  features (per pair, from the ACGT strings):
    x0 = log2(p+1)/16, x1 = log2(t+1)/16, x2 = log2(|p-t|+1)/16,
    x3 = found/16 where, for i = 0..15, the pattern 12-mer at
         pos_i = ((2i+1)(p-12)) // 32 counts as found when it occurs in the text
         at some j in [max(0, q_i - W), min(t-12, q_i + W)],
         q_i = pos_i (t-12) // max(p-12, 1), W = 96 + |p-t|   (0 if p or t < 12)
  network: 4 -> 64 -> 64 -> 1, ReLU, weights rounded to BF16 (round to nearest,
    ties to even) and every product accumulated in FP32; loss = mean squared error.

The gradient restatement is pinned by central finite differences in float64
(tests/test_costmodel.py), the features by brute force on crafted pairs.
"""
import numpy as np

K_IN, K_H = 4, 64
W1 = 0
B1 = W1 + K_H * K_IN
W2 = B1 + K_H
B2 = W2 + K_H * K_H
W3 = B2 + K_H
B3 = W3 + K_H
N_PARAMS = B3 + 1
KMER = 12


def bf16_round(a):
    """FP32 -> BF16 (round to nearest, ties to even) -> FP32; inf/nan unchanged."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) & 0xFFFF0000).astype(np.uint32)
    r = np.where(special, u.astype(np.uint32), r)
    return r.view(np.float32).reshape(a.shape)


def unpack(w):
    w = np.asarray(w)
    return (w[W1:B1].reshape(K_H, K_IN), w[B1:W2], w[W2:B2].reshape(K_H, K_H), w[B2:W3],
            w[W3:B3], w[B3])


def features_one(p, t):
    """The four features of one pair (ASCII strings)."""
    lp, lt = len(p), len(t)
    x = np.zeros(4, np.float32)
    x[0] = np.float32(np.log2(np.float32(lp) + np.float32(1))) / np.float32(16)
    x[1] = np.float32(np.log2(np.float32(lt) + np.float32(1))) / np.float32(16)
    x[2] = np.float32(np.log2(np.float32(abs(lp - lt)) + np.float32(1))) / np.float32(16)
    if lp < KMER or lt < KMER:
        return x
    w = 96 + abs(lp - lt)
    found = 0
    for i in range(16):
        pos = ((2 * i + 1) * (lp - KMER)) // 32
        key = p[pos:pos + KMER]
        q = (pos * (lt - KMER)) // max(lp - KMER, 1)
        lo, hi = max(0, q - w), min(lt - KMER, q + w)
        if lo <= hi and t.find(key, lo, hi + KMER) >= 0:
            found += 1
    x[3] = np.float32(found) / np.float32(16)
    return x


def features(pairs):
    return np.stack([features_one(p, t) for p, t in pairs]) if pairs else np.zeros((0, 4), np.float32)


def forward(w_bf16, X, dtype=np.float32):
    """Returns (y, cache) for samples X[n,4] with (already rounded) weights."""
    w1, b1, w2, b2, w3, b3 = [np.asarray(a, dtype) for a in unpack(w_bf16)]
    X = np.asarray(X, dtype)
    z1 = X @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    return y, (X, z1, h1, z2, h2)


def loss(w_bf16, X, y_true, dtype=np.float32):
    y, _ = forward(w_bf16, X, dtype)
    return np.mean((y - np.asarray(y_true, dtype)) ** 2)


def gradient(w_bf16, X, y_true, dtype=np.float32):
    """d(mean squared error)/d(weights), in the parameter layout."""
    w1, b1, w2, b2, w3, b3 = [np.asarray(a, dtype) for a in unpack(w_bf16)]
    y, (X, z1, h1, z2, h2) = forward(w_bf16, X, dtype)
    n = len(X)
    g = (2.0 / n) * (y - np.asarray(y_true, dtype))          # dL/dy
    d2 = np.outer(g, w3) * (z2 > 0)                          # dL/dz2
    d1 = (d2 @ w2) * (z1 > 0)                                # dL/dz1
    out = np.zeros(N_PARAMS, dtype)
    out[W1:B1] = (d1.T @ X).reshape(-1)
    out[B1:W2] = d1.sum(0)
    out[W2:B2] = (d2.T @ h1).reshape(-1)
    out[B2:W3] = d2.sum(0)
    out[W3:B3] = h2.T @ g
    out[B3] = g.sum()
    return out


def score_to_target(score):
    return np.log2(1.0 + np.maximum(np.asarray(score, np.float64), 0)) / 16.0


def target_to_score(y):
    return np.exp2(16.0 * np.asarray(y, np.float64)) - 1.0
