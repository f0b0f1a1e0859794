"""CPU pins of the warm-start network oracle (oracle/warmstart_oracle.py)."""

import numpy as np
import pytest
import torch

from oracle import warmstart_oracle as W


def test_param_count():
    # 3x3x3 convolutions on levels of 16/32/64/128 channels with skip
    # concatenation: 1 464 480 weights + 710 biases (6-channel head)
    assert W.param_count() == 1464480 + 710
    assert 1.4e6 < W.param_count() < 1.5e6


def test_periodic_conv_matches_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 4, 5, 3))
    w = rng.standard_normal((4, 27, 3))
    b = rng.standard_normal(4)
    ref = W.conv_periodic_loop(x, w, b)
    got = W.conv_periodic(torch.tensor(x), torch.tensor(w), torch.tensor(b)).numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_periodic_conv_single_tap():
    """A weight with one nonzero tap shifts the input: out[i] = in[i + d] (periodic)."""
    x = np.arange(2 * 4 * 4 * 4, dtype=np.float64).reshape(1, 2, 4, 4, 4)
    for tap in (0, 5, 13, 22, 26):
        dx, dy, dz = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        w = np.zeros((2, 27, 2))
        w[0, tap, 1] = 1.0
        w[1, tap, 0] = 2.0
        out = W.conv_periodic(torch.tensor(x), torch.tensor(w), torch.zeros(2, dtype=torch.float64)).numpy()
        for i in range(4):
            for j in range(4):
                for k in range(4):
                    src = ((i + dx) % 4, (j + dy) % 4, (k + dz) % 4)
                    assert out[0, 0, i, j, k] == x[0, 1][src]
                    assert out[0, 1, i, j, k] == 2.0 * x[0, 0][src]


def test_pool_and_upsample_loops():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 2, 4, 6, 2))
    p = W.pool(torch.tensor(x)).numpy()
    for i in range(2):
        for j in range(3):
            assert np.isclose(p[0, 1, i, j, 0], x[0, 1, 2 * i:2 * i + 2, 2 * j:2 * j + 2, 0:2].mean())
    u = W.upsample(torch.tensor(p)).numpy()
    assert u.shape == (1, 2, 4, 6, 2)
    assert u[0, 0, 3, 5, 1] == p[0, 0, 1, 2, 0]


def test_bf16_rounding_is_nearest_even():
    v = torch.tensor([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7 + 2.0 ** -9], dtype=torch.float64)
    r = W.bf16_round(v).numpy()
    # ties to even: 1 + 2^-8 -> 1; 1 + 3 2^-8 -> 1 + 2^-6; above the tie rounds up
    assert r[0] == 1.0 and r[1] == 1.0 + 2.0 ** -6 and r[2] == 1.0 + 2.0 ** -7


def _tiny_params(rng, scale=0.3):
    return {name: (rng.uniform(-scale, scale, (co, 27, ci)) / np.sqrt(ci), rng.uniform(-0.1, 0.1, co))
            for name, ci, co, _ in W.LAYERS}


def test_autograd_matches_finite_differences():
    """Gradients of the FP64 oracle (no rounding) against central differences."""
    rng = np.random.default_rng(2)
    params = _tiny_params(rng)
    lam = rng.uniform(2.5, 20.0, (1, 8, 8, 8))
    mu = rng.uniform(1.1, 29.0, (1, 8, 8, 8))
    eb = np.eye(6)[:2]
    norm = (1.9, 1.0, 1.7, 0.6)
    target = rng.standard_normal((2, 6, 8, 8, 8)) * 0.1
    loss, _, grads = W.loss_and_grads(params, lam, mu, eb, 2, norm, target, bf16=False)
    for name, idx in (("enc1a", (3, 13, 1)), ("botb", (7, 4, 100)), ("dec1a", (2, 20, 40)), ("head", (5, 26, 9))):
        h = 1e-5
        pp = {k: (w.copy(), b.copy()) for k, (w, b) in params.items()}
        pp[name][0][idx] += h
        lp, _, _ = W.loss_and_grads(pp, lam, mu, eb, 2, norm, target, bf16=False)
        pp[name][0][idx] -= 2 * h
        lm, _, _ = W.loss_and_grads(pp, lam, mu, eb, 2, norm, target, bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grads[name][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, fd, grads[name][0][idx])
