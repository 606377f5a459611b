"""Pinning of the cost model's CPU oracle (oracle/mlp.c).

The model has no LELA counterpart, so the oracle is pinned by mathematics:
central finite differences of the loss (with parameters and steps that are
exact in float32, so the only error is the difference quotient's own),
the closed form of a network whose hidden layers are zero, and the first
Adam step (m-hat = g, v-hat = g^2, so each parameter moves by lr * sign(g)).
A dropped ReLU mask, a transposed weight block or a wrong loss scale fails
the finite-difference check."""
import numpy as np
import pytest

import oracle_lib as orc

NP = orc.MLP_PARAMS
OB3 = 1376


def exact_params(rng, scale=1.0):
    # multiples of 2^-12 in [-scale, scale]: w + 2^-20 is exact in float32
    return (np.round(rng.uniform(-scale, scale, NP) * 4096) / 4096).astype(np.float32)


def data(rng, n):
    x = (np.round(rng.uniform(-1, 1, (n, 5)) * 4096) / 4096).astype(np.float32)
    t = rng.uniform(-3, 3, (n, 4)).astype(np.float32)
    mask = (rng.random((n, 4)) < 0.7).astype(np.uint8)
    mask[0, :] = 1
    return x, t, mask


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    w = exact_params(rng, 0.5)
    x, t, mask = data(rng, 9)
    loss, g = orc.mlp_loss_grad(w, x, t, mask)
    h = 2.0 ** -20
    # every bias and a spread of weights of each layer
    idx = list(range(0, NP, 7)) + list(range(160, 192)) + list(range(1216, 1248)) + list(range(1376, 1380))
    bad = []
    for q in sorted(set(idx)):
        wp, wm = w.copy(), w.copy()
        wp[q] += np.float32(h)
        wm[q] -= np.float32(h)
        assert float(wp[q]) - float(w[q]) == h and float(w[q]) - float(wm[q]) == h
        lp, _ = orc.mlp_loss_grad(wp, x, t, mask)
        lm, _ = orc.mlp_loss_grad(wm, x, t, mask)
        fd = (lp - lm) / (2 * h)
        if abs(fd - g[q]) > 1e-6 * max(1.0, abs(g[q])):
            bad.append((q, fd, g[q]))
    assert not bad, bad[:5]
    assert loss > 0


def test_mlp_zero_hidden_layers_closed_form():
    rng = np.random.default_rng(3)
    w = np.zeros(NP, dtype=np.float32)
    w[OB3:] = [0.5, -1.0, 2.0, 0.25]
    x, t, mask = data(rng, 13)
    y = orc.mlp_forward(w, x)
    assert np.array_equal(y, np.tile(w[OB3:].astype(np.float64), (13, 1)))
    loss, g = orc.mlp_loss_grad(w, x, t, mask)
    M = mask.sum()
    d = (w[OB3:][None, :].astype(np.float64) - t) * mask
    assert loss == pytest.approx((d * d).sum() / M, rel=1e-12)
    assert np.allclose(g[OB3:], 2 * d.sum(axis=0) / M, rtol=1e-12)
    assert not np.any(g[:OB3])      # zero hidden activations: no weight gradient


def test_mlp_adam_first_step_moves_by_lr_sign():
    rng = np.random.default_rng(5)
    w = exact_params(rng, 0.5)
    x, t, mask = data(rng, 17)
    loss0, g = orc.mlp_loss_grad(w, x, t, mask)
    lr = 1e-3
    w1, hist = orc.mlp_adam(w, x, t, mask, 1, lr)
    assert hist[0] == pytest.approx(loss0, rel=1e-12)
    big = np.abs(g) > 1e-4
    step = w1.astype(np.float64) - w.astype(np.float64)
    assert np.allclose(step[big], -lr * np.sign(g[big]), rtol=1e-3, atol=1e-7)


def test_mlp_adam_decreases_loss():
    rng = np.random.default_rng(9)
    w = exact_params(rng, 0.3)
    x, t, mask = data(rng, 64)
    _, hist = orc.mlp_adam(w, x, t, mask, 200, 1e-2)
    assert hist[-1] < 0.5 * hist[0]
