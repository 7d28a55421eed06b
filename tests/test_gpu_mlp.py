"""Cost-model kernels (pi_compute_amd/csrc/mlp.hip) against the double-precision
oracle (oracle/mlp_oracle.c, itself pinned by finite differences in
test_mlp_oracle.py).  Tolerances are for FP32 against FP64: forward 1e-5
relative to the output scale, gradient 2e-4 relative to the largest entry."""
import numpy as np
import pytest

import pi_compute_amd as pc
from mlp_data import params, samples

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,n", [(2, 1), (2, 257), (3, 5000), (8, 100)])
def test_forward_matches_oracle(oracle, engine, k, n):
    w = params(k, n)
    x, _, _ = samples(n, k, n)
    y = pc.mlp_forward(w.astype(np.float32), k, x.astype(np.float32))
    want = oracle.mlp_forward(w.astype(np.float32).astype(np.float64), k, x.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(y - want)) <= 1e-5 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("k,n,seed", [(2, 1, 1), (2, 300, 2), (2, 20000, 3), (5, 4097, 4)])
def test_gradient_matches_oracle(oracle, engine, k, n, seed):
    w = params(k, seed).astype(np.float32)
    x, path, t = samples(n, k, seed)
    x, t = x.astype(np.float32), t.astype(np.float32)
    g, loss = pc.mlp_gradient(w, k, x, path, t)
    want, want_loss = oracle.mlp_gradient(w.astype(np.float64), k, x.astype(np.float64), path, t.astype(np.float64))
    scale = np.max(np.abs(want))
    assert np.max(np.abs(g - want)) <= 2e-4 * scale, np.argmax(np.abs(g - want))
    assert loss == pytest.approx(want_loss, rel=1e-4)


def test_gradient_is_deterministic(engine):
    w = params(2, 7).astype(np.float32)
    x, path, t = samples(10000, 2, 7)
    g1, l1 = pc.mlp_gradient(w, 2, x, path, t)
    g2, l2 = pc.mlp_gradient(w, 2, x, path, t)
    assert np.array_equal(g1, g2) and l1 == l2


def test_training_follows_oracle_adam(oracle, engine):
    """The first Adam steps on the device track the oracle's steps in FP64."""
    k, steps, lr = 2, 5, 1e-3
    w0 = params(k, 21).astype(np.float32)
    x, path, t = samples(2000, k, 21)
    x, t = x.astype(np.float32), t.astype(np.float32)
    w_dev, hist = pc.mlp_train(w0, k, x, path, t, steps, lr)
    w = w0.astype(np.float64)
    m, v = np.zeros_like(w), np.zeros_like(w)
    losses = []
    for s in range(1, steps + 1):
        g, L = oracle.mlp_gradient(w, k, x.astype(np.float64), path, t.astype(np.float64))
        losses.append(L)
        oracle.mlp_adam(w, g, m, v, s, lr)
    assert np.allclose(hist, losses, rtol=1e-4)
    # Adam moves each parameter by about lr per step: the paths agree to a small fraction of that
    assert np.max(np.abs(w_dev - w)) < 0.05 * lr * steps


def test_training_fits_a_cost_surface(engine):
    """Trained on a synthetic two-path cost surface, the network's choices match
    the cheaper path far better than a single threshold on operand size."""
    rng = np.random.default_rng(5)
    n = 8000
    lb = rng.uniform(4, 22, n)  # log2 limbs (larger)
    ls = lb - rng.uniform(0, 6, n)
    lbat = rng.integers(0, 12, n).astype(np.float64)
    sq = rng.integers(0, 2, n).astype(np.float64)
    cpu = 1.45 * np.log(2.0) * (lb + ls) / 2 - 9 - 0.2 * sq
    gpu = np.log(0.05 + np.exp(0.8 * np.log(2.0) * (lb + ls) / 2 - 8) * 2.0 ** -lbat) - 0.1 * sq
    noise = rng.uniform(0, 1, (n, 4))  # four inputs the surface does not depend on
    x = np.concatenate([np.stack([lb, ls, lbat, sq], 1), noise], 1)
    mean, sd = x.mean(0), x.std(0)
    xs = ((x - mean) / sd).astype(np.float32)
    path = rng.integers(0, 2, n).astype(np.int32)
    t = np.where(path == 0, cpu, gpu).astype(np.float32)
    w0 = params(2, 3).astype(np.float32)
    w, hist = pc.mlp_train(w0, 2, xs, path, t, 3000, 3e-3)
    assert hist[-1] < 0.05 * hist[0]
    y = pc.mlp_forward(w, 2, xs)
    best = gpu < cpu
    choice = y[:, 1] < y[:, 0]
    thr = ls >= np.log2(10000)
    assert np.mean(choice == best) > 0.95
    assert np.mean(choice == best) > np.mean(thr == best) + 0.1
