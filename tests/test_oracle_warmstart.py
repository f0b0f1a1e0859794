"""CPU pins of the warm-start network oracle (oracle/warmstart_oracle.py)."""

import numpy as np
import pytest
import torch

from oracle import warmstart_oracle as W


def test_param_count():
    # 3x3x3 convolutions on levels of 16/32/64/128 channels with skip
    # concatenation: 1 469 664 weights + 722 biases (18-channel displacement head)
    assert W.param_count() == 1469664 + 722
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
    lam = rng.uniform(2.5, 20.0, (2, 8, 8, 8))
    mu = rng.uniform(1.1, 29.0, (2, 8, 8, 8))
    norm = (1.9, 1.0, 1.7, 0.6)
    ext = (8.0, 6.0, 7.0)
    target = rng.standard_normal((2, 36, 8, 8, 8)) * 0.1
    loss, _, grads = W.loss_and_grads(params, lam, mu, norm, target, ext, wm=0.3, wr=1.0, bf16=False)
    for name, idx in (("enc1a", (3, 13, 1)), ("botb", (7, 4, 100)), ("dec1a", (2, 20, 40)), ("head", (15, 26, 9))):
        h = 1e-5
        pp = {k: (w.copy(), b.copy()) for k, (w, b) in params.items()}
        pp[name][0][idx] += h
        lp, _, _ = W.loss_and_grads(pp, lam, mu, norm, target, ext, wm=0.3, wr=1.0, bf16=False)
        pp[name][0][idx] -= 2 * h
        lm, _, _ = W.loss_and_grads(pp, lam, mu, norm, target, ext, wm=0.3, wr=1.0, bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grads[name][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, fd, grads[name][0][idx])


def test_divergence_matches_the_stencil_oracle():
    """The torch divergence against ls_oracle.backward_divergence (pinned by the
    explicit stencil loop in tests/test_oracle.py)."""
    from oracle import ls_oracle as O

    rng = np.random.default_rng(4)
    sig = rng.standard_normal((6, 5, 4, 6))
    gs = O.GridSpec(5, 4, 6, 1.3, 0.7, 0.9)
    ref = O.backward_divergence(sig, gs)
    got = W.divergence(torch.tensor(sig), (1.3, 0.7, 0.9)).numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_residual_term_is_the_solvers_relative_residual():
    """With an equilibrium target (the FP64 oracle's solution), the residual
    term of one load case is (R(e_k + pred) / R(e_k))^2 times the squared ratio
    of the mean-stress norms that normalise the solver's R (Parseval: the
    backward divergence has the magnitude |xi~ . sigma_hat| per mode)."""
    from oracle import ls_oracle as O

    n = 8
    gs = O.GridSpec(n, n, n, 8.0, 8.0, 8.0)
    rng = np.random.default_rng(9)
    fib = rng.random((n, n, n)) < 0.4
    lam = np.where(fib, 19.44, 2.593)
    mu = np.where(fib, 29.17, 1.111)
    target = np.zeros((1, 36, n, n, n))
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1.0
        target[0, 6 * k:6 * k + 6] = O.lippmann_schwinger_rfft(lam, mu, e, gs, rtol=1e-13, depth=4,
                                                                 maxiter=5000)[0] - e[:, None, None, None]
    pred = target + 0.05 * rng.standard_normal(target.shape)
    sp = O._RealSpectral(gs)
    T = lambda a: torch.tensor(a)  # noqa: E731
    for k in (0, 4):
        # the other load cases predicted exactly: only load case k contributes
        pk = target.copy()
        pk[0, 6 * k:6 * k + 6] = pred[0, 6 * k:6 * k + 6]
        term = float(W.residual_loss(T(pk), T(target), T(lam[None]), T(mu[None]), (8.0, 8.0, 8.0), 0.0, 1.0)) * 6
        e = np.zeros(6)
        e[k] = 1.0
        eps = e[:, None, None, None] + pred[0, 6 * k:6 * k + 6]
        R, _ = sp.residual(sp.fwd(O.compute_sigma(lam, mu, eps.copy())))
        R0, _ = sp.residual(sp.fwd(O.compute_sigma(lam, mu, np.broadcast_to(e[:, None, None, None],
                                                                          (6, n, n, n)).copy())))
        m_ws = O.compute_sigma(lam, mu, eps.copy()).mean(axis=(1, 2, 3))
        m_0 = O.compute_sigma(lam, mu, np.broadcast_to(e[:, None, None, None], (6, n, n, n)).copy()).mean(axis=(1, 2, 3))
        nv = lambda m: np.sqrt(np.sum(O.VOIGT_WEIGHTS * m ** 2))  # noqa: E731
        want = (R * nv(m_ws) / (R0 * nv(m_0))) ** 2
        assert abs(term - want) <= 1e-6 * want, (k, term, want)


def test_symgrad_is_the_adjoint_of_the_divergence_and_compatible():
    """D^+ = -(D^-)^T (inner-product identity against the stencil oracle's
    backward derivative), and sym(D^+ u) is unchanged by the oracle's
    projection onto compatible zero-mean fields (Gamma0 with C0 = identity),
    i.e. the solver's projection of such a start is the identity."""
    from oracle import ls_oracle as O

    rng = np.random.default_rng(12)
    shape, ext = (6, 5, 4), (1.3, 0.7, 0.9)
    gs = O.GridSpec(*shape, *ext)
    u = rng.standard_normal((1, 18) + shape)
    eps = W.symgrad(torch.tensor(u), ext).numpy()[0]
    for k in (0, 5):
        ek = eps[6 * k:6 * k + 6]
        h = rng.standard_normal((6,) + shape)
        # <sym(D+ u), h>_w = -<u, D^- . h> (Voigt weights on the shear)
        lhs = float(np.einsum("aijk,aijk,a->", ek, h, O.VOIGT_WEIGHTS))
        rhs = -float(np.sum(u[0, 3 * k:3 * k + 3] * O.backward_divergence(h, gs)))
        assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))
        sp = O._RealSpectral(gs)
        proj = sp.inv(-O.fourier_solve(sp.fwd(ek), 0.0, 0.5, sp.xizero))
        assert np.abs(proj - ek).max() <= 1e-10 * np.abs(ek).max()
