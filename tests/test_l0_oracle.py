"""Pins of the learned-l0 oracle (oracle/l0_oracle.py) on the CPU."""
import math

import numpy as np
import pytest

from oracle import l0_oracle as lo
from oracle import zolo_oracle as zo


def _data(rng, N=40):
    F = rng.standard_normal((N, lo.N_FEAT))
    t = rng.standard_normal(N)
    P = lo.init_params(rng, mu=rng.standard_normal(lo.N_FEAT) * 0.1, sd=1.0 + rng.random(lo.N_FEAT))
    return P, F, t


def test_sizes():
    assert lo.N_TRAIN == 5313 and lo.N_PARAM == 5346


@pytest.mark.parametrize("tau", [0.05, 0.5])
def test_gradient_matches_finite_differences(tau):
    rng = np.random.default_rng(0)
    P, F, t = _data(rng)
    loss, g = lo.loss_grad(P, F, t, tau)
    assert loss == pytest.approx(lo.pinball(t, lo.forward(P, F)[0], tau))
    o = 2 * lo.N_FEAT
    idx = rng.choice(lo.N_TRAIN, size=60, replace=False)
    idx = np.concatenate([idx, [0, lo.N_TRAIN - 1, lo.N_HID * lo.N_FEAT, lo.N_TRAIN - 2]])
    h = 1e-6
    for k in idx:
        Pp = P.copy(); Pp[o + k] += h
        Pm = P.copy(); Pm[o + k] -= h
        fd = (lo.pinball(t, lo.forward(Pp, F)[0], tau) - lo.pinball(t, lo.forward(Pm, F)[0], tau)) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9), k


def test_gradient_mutation_is_caught():
    """A dropped tanh derivative in the first layer fails the finite-difference check."""
    rng = np.random.default_rng(1)
    P, F, t = _data(rng)
    mu, sd, W1, b1, W2, b2, w3, b3, _ = lo.unpack(P)
    y, (Xn, H1, H2) = lo.forward(P, F)
    d3 = ((y > t).astype(float) - 0.3) / len(y)
    d2 = np.outer(d3, w3) * (1 - H2 * H2)
    d1_wrong = d2 @ W2
    g_wrong = (d1_wrong.T @ Xn).ravel()
    _, g = lo.loss_grad(P, F, t, 0.3)
    assert not np.allclose(g_wrong, g[:lo.N_HID * lo.N_FEAT], rtol=1e-3)


def test_pinball_definition():
    assert lo.pinball([1.0, 2.0], [0.0, 3.0], 0.25) == pytest.approx((0.25 * 1 + 0.75 * 1) / 2)
    assert lo.pinball([0.0], [0.0], 0.1) == 0.0


def test_adam_first_step_moves_by_lr_sign():
    rng = np.random.default_rng(2)
    P, F, t = _data(rng)
    _, g = lo.loss_grad(P, F, t, 0.1)
    P1, m, v = lo.adam_step(P, g, np.zeros_like(g), np.zeros_like(g), 1e-3, 1)
    d = (P1 - P)[2 * lo.N_FEAT:2 * lo.N_FEAT + lo.N_TRAIN]
    nz = np.abs(g) > 1e-4   # eps = 1e-8 changes the step by eps / |g| relative
    assert nz.sum() > lo.N_TRAIN // 2
    assert np.allclose(d[nz], -1e-3 * np.sign(g[nz]), rtol=2e-4, atol=0)
    assert np.array_equal(P1[:2 * lo.N_FEAT], P[:2 * lo.N_FEAT]) and P1[-1] == P[-1]


def test_lanczos_projection_and_interlacing():
    rng = np.random.default_rng(3)
    B = rng.standard_normal((50, 50))
    M = B.T @ B / 50
    a, b, Q = lo.lanczos(M, lo.start_vector(50))
    assert len(a) == lo.N_LANCZOS
    T = Q.T @ M @ Q
    k = len(a)
    Tt = np.diag(a) + np.diag(b[:k - 1], 1) + np.diag(b[:k - 1], -1)
    assert np.allclose(T, Tt, atol=1e-12)
    assert np.allclose(Q.T @ Q, np.eye(k), atol=1e-13)
    th = lo.ritz_values(a, b)
    lam = np.sort(np.linalg.eigvalsh(M))[::-1]
    assert np.all(th <= lam[:k] + 1e-12) and np.all(th >= lam[-k:] - 1e-12)
    assert th[0] == pytest.approx(lam[0], rel=2e-3)


def test_features_closed_form_orthogonal():
    rng = np.random.default_rng(4)
    n = 40
    Q = zo.haar(n, rng)
    f, smax = lo.features(3.0 * Q)
    # X = Q / sqrt(n): X^T X = I / n, so every Ritz value is 1/n and n theta_1 = 1
    assert np.allclose(f[1:9], math.log10(1.0 / n), atol=1e-12)
    assert f[15] == pytest.approx(1.0, rel=1e-12)
    assert smax == pytest.approx(1.0 / math.sqrt(n), rel=1e-12)
    assert f[0] == pytest.approx(math.log2(n))
    # sketch of an orthogonal X: R = chol of Omega^T Omega / n
    Om = lo.sketch_matrix(n)
    R = np.linalg.qr(Om / math.sqrt(n))[1]
    assert np.allclose(f[11:15], np.log10(np.abs(np.diag(R))), atol=1e-12)


def test_features_closed_form_diagonal():
    d = np.array([3.0, -1.0, 0.5, 2.0, 0.25, 1.5])
    f, smax = lo.features(np.diag(d))
    fro = np.linalg.norm(d)
    th = np.sort(d * d / fro ** 2)[::-1]
    # n <= 8: the Krylov space is the whole space, the Ritz values are exact
    assert np.allclose(f[1:7], np.log10(th), atol=1e-12)
    assert np.allclose(f[7:9], math.log10(th[-1]), atol=1e-12)
    assert f[9] == pytest.approx(math.log10(3.0 / fro)) and f[10] == pytest.approx(math.log10(3.0 / fro))
    assert smax == pytest.approx(3.0 / fro, rel=1e-12)


def test_beta_of_margin():
    P = np.zeros(lo.N_PARAM)
    P[-1] = 0.5
    assert lo.beta_of(P, np.array([-3.0, 1.0]), np.array([0.2, 0.2])) == pytest.approx([0.2 * 10 ** -3.5, 0.2])
