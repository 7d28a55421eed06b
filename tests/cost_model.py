"""The dispatch cost model read and evaluated in numpy (test helper).

File format and forward pass of pi_compute_amd/host/mlp_dispatch.cpp
(MlpModel::load, MlpModel::features, MlpModel::forward): eight features
standardised by the model's mean and scale (the batch feature capped at its
trained maximum), 8 -> 32 ReLU -> 32 ReLU -> K outputs, ln(ms) per path;
parameter layout of include/pintt_mlp.h: W1 (32x8), b1, W2 (32x32), b2,
W3 (Kx32), b3.

Features: log2 larger limbs, log2 smaller limbs, log2 batch, square, and for
m = 1 .. 4 the padding log2(2^k / (2 h + 2 bn - 1)) of the transform (length
2^k, chunks of h limbs) of the product cut into m chunks, -1 where that cut is
not allowed (squares are never cut).  The cut rule is restated here
(chunk_plan) and checked against libpintt's gpu_ntt_chunk_plan in the tests.
"""
import math

import numpy as np

H, IN = 32, 8
MAX_LOG2N = 25


def _log2n(x, y):
    """The engine's transform length: the least 2^k >= 2x + 2y - 1, 10 <= k <= 25."""
    need, k = 2 * x + 2 * y - 1, 10
    while (1 << k) < need:
        k += 1
    return k if k <= MAX_LOG2N else -1


def chunk_plan(an, bn, m):
    """(chunk limbs, log2 transform length) of an an x bn product cut into m
    chunks, or None: chunks of ceil(an / m) limbs rounded up to 16, none
    shorter than b, no empty last chunk, a shorter transform than uncut."""
    if an < bn:
        an, bn = bn, an
    full = _log2n(an, bn)
    if full < 0 or not 1 <= m <= 4:
        return None
    if m == 1:
        return an, full
    h = -(-an // m)
    h = (h + 15) & ~15
    if h < bn or (m - 1) * h >= an:
        return None
    k = _log2n(h, bn)
    return (h, k) if 0 <= k < full else None


def load_model(path):
    tok = open(path).read().split()
    assert tok[0] == "pintt-mlp" and tok[1] == "2" and tok[2] == "outputs"
    k = int(tok[3])
    pos, arrs = 4, {}
    for name in ("mean", "scale", "lo", "hi"):
        assert tok[pos] == name
        arrs[name] = np.array([float(x) for x in tok[pos + 1:pos + 1 + IN]], np.float32)
        pos += 1 + IN
    assert tok[pos] == "params"
    n = int(tok[pos + 1])
    w = np.array([float(x) for x in tok[pos + 2:pos + 2 + n]], np.float32)
    assert w.size == H * IN + H + H * H + H + (H + 1) * k
    return {"k": k, "w": w, **arrs}


def features(big, small, batch, square):
    f = [math.log2(big), math.log2(small), math.log2(max(batch, 1)), 1.0 if square else 0.0]
    for m in range(1, 5):
        c = chunk_plan(big, small, m) if (m == 1 or not square) else None
        f.append(c[1] - math.log2(2 * c[0] + 2 * small - 1) if c else -1.0)
    return np.array(f, np.float32)


def in_range(m, f):
    return all(m["lo"][j] <= f[j] <= m["hi"][j] for j in (0, 1)) and f[2] >= m["lo"][2]


def forward(m, f):
    x = f.copy()
    x[2] = min(x[2], m["hi"][2])
    x = (x - m["mean"]) / m["scale"]
    w, k = m["w"], m["k"]
    o = 0
    w1 = w[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = w[o:o + H]; o += H
    w2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    w3 = w[o:o + k * H].reshape(k, H); o += k * H
    b3 = w[o:o + k]
    h1 = np.maximum(w1 @ x + b1, 0)
    h2 = np.maximum(w2 @ h1 + b2, 0)
    return w3 @ h2 + b3
