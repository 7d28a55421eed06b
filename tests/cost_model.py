"""The dispatch cost model read and evaluated in numpy (test helper).

File format and forward pass of pi_compute_amd/host/mlp_dispatch.cpp
(MlpModel::load, MlpModel::forward): features (log2 larger limbs, log2 smaller
limbs, log2 batch, square) standardised by the model's mean and scale (the
batch feature capped at its trained maximum), 4 -> 32 ReLU -> 32 ReLU -> K
outputs, ln(ms) per path; parameter layout of include/pintt_mlp.h: W1 (32x4),
b1, W2 (32x32), b2, W3 (Kx32), b3.
"""
import math

import numpy as np

H, IN = 32, 4


def load_model(path):
    tok = open(path).read().split()
    assert tok[0] == "pintt-mlp" and tok[1] == "1" and tok[2] == "outputs"
    k = int(tok[3])
    pos, arrs = 4, {}
    for name in ("mean", "scale", "lo", "hi"):
        assert tok[pos] == name
        arrs[name] = np.array([float(x) for x in tok[pos + 1:pos + 5]], np.float32)
        pos += 5
    assert tok[pos] == "params"
    n = int(tok[pos + 1])
    w = np.array([float(x) for x in tok[pos + 2:pos + 2 + n]], np.float32)
    assert w.size == H * IN + H + H * H + H + (H + 1) * k
    return {"k": k, "w": w, **arrs}


def features(big, small, batch, square):
    return np.array([math.log2(big), math.log2(small), math.log2(max(batch, 1)), 1.0 if square else 0.0],
                    np.float32)


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
