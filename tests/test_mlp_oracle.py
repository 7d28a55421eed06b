"""Pins the cost model's CPU oracle (oracle/mlp_oracle.c) before the HIP
kernels are checked against it: the gradient against central finite
differences of the loss, the forward pass against a numpy evaluation of the
layout stated in include/pintt_mlp.h, Adam against its first-step closed form,
and a planted mistake."""
import numpy as np
import pytest

from mlp_data import HEAD, H, IN, param_count, params, samples


def _np_forward(w, k, x):
    o = 0
    W1 = w[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    W3 = w[o:o + k * H].reshape(k, H); o += k * H
    b3 = w[o:o + k]
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


@pytest.mark.parametrize("k", [1, 2, 5, 8])
def test_param_count(oracle, k):
    assert oracle.mlp_param_count(k) == param_count(k) == HEAD + 33 * k


@pytest.mark.parametrize("k", [2, 3])
def test_forward_matches_layout(oracle, k):
    w = params(k, 1)
    x, _, _ = samples(64, k, 1)
    assert np.allclose(oracle.mlp_forward(w, k, x), _np_forward(w, k, x), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k,seed", [(2, 3), (4, 4)])
def test_gradient_matches_finite_differences(oracle, k, seed):
    w = params(k, seed)
    x, path, t = samples(40, k, seed)
    g, L = oracle.mlp_gradient(w, k, x, path, t)
    assert L == pytest.approx(oracle.mlp_loss(w, k, x, path, t), rel=1e-14)
    eps = 1e-6
    rng = np.random.default_rng(seed)
    # every bias and output weight, and a sample of the hidden weights
    idx = list(range(H * IN, H * IN + H)) + list(range(HEAD, w.size)) + list(rng.choice(HEAD, 200, replace=False))
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (oracle.mlp_loss(wp, k, x, path, t) - oracle.mlp_loss(wm, k, x, path, t)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_loss_closed_form_with_zero_weights(oracle):
    k = 2
    w = np.zeros(param_count(k))
    w[HEAD + 32 * k:] = [0.5, -1.0]
    x, path, t = samples(30, k, 9)
    pred = np.where(path == 0, 0.5, -1.0)
    g, L = oracle.mlp_gradient(w, k, x, path, t)
    assert L == pytest.approx(0.5 * np.mean((pred - t) ** 2))
    # only output biases move: dL/db3[j] = mean over samples of path j of the error, scaled by 1/n
    for j in range(k):
        assert g[HEAD + 32 * k + j] == pytest.approx(np.sum((pred - t)[path == j]) / len(t))
    assert np.all(g[:HEAD + 32 * k] == 0)


def test_adam_first_step_closed_form(oracle):
    # After one step from m = v = 0 each parameter moves by lr * sign(g) (up to eps).
    w = np.linspace(-1, 1, 50)
    g = np.linspace(-3, 2, 50) + 0.01
    m, v = np.zeros(50), np.zeros(50)
    w0 = w.copy()
    oracle.mlp_adam(w, g, m, v, 1, 0.01)
    assert np.allclose(w, w0 - 0.01 * np.sign(g), atol=1e-7)


def test_planted_mistake_is_caught(oracle):
    """A gradient with one entry off is rejected by the finite-difference check."""
    k = 2
    w = params(k, 11)
    x, path, t = samples(20, k, 11)
    g, _ = oracle.mlp_gradient(w, k, x, path, t)
    i = HEAD + 5
    g[i] *= 1.01
    wp, wm = w.copy(), w.copy()
    wp[i] += 1e-6
    wm[i] -= 1e-6
    fd = (oracle.mlp_loss(wp, k, x, path, t) - oracle.mlp_loss(wm, k, x, path, t)) / 2e-6
    assert abs(fd - g[i]) > 1e-6 + 1e-5 * abs(g[i])
