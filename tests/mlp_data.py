"""Seeded parameters and samples for the cost-model tests."""
import numpy as np

IN, H = 8, 32
HEAD = H * IN + H + H * H + H  # parameters before the output layer (W3, b3)


def param_count(k):
    return HEAD + (H + 1) * k


def params(k, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(param_count(k))
    o = 0
    for rows, cols in ((H, IN), (H, H), (k, H)):
        lim = np.sqrt(6.0 / cols)
        w[o:o + rows * cols] = rng.uniform(-lim, lim, rows * cols)
        o += rows * cols
        w[o:o + rows] = rng.uniform(-0.1, 0.1, rows)
        o += rows
    return w


def samples(n, k, seed):
    rng = np.random.default_rng(seed + 1000)
    x = rng.normal(size=(n, IN))
    x[:, 3] = rng.integers(0, 2, n)
    path = rng.integers(0, k, n).astype(np.int32)
    t = rng.normal(size=n)
    return x, path, t
