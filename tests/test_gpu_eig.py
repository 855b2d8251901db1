"""Spectral division on the GPU: PQDWHSUBIT (pqdwhsubit.f90:2), ZoloEig
(ZoloEig.f90:2) and the driver of pqdwheig.f90:113-144 against LAPACK dsyevd
(numpy.linalg.eigvalsh / eigh) and the oracle's restatement: the number of
eigenvalues below sigma, those eigenvalues, the residual ||A V - V diag(w)||,
V^T V = I and the largest principal angle between span(V) and the exact
invariant subspace."""
import math

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu

TOL1 = 10.0 * zo.DLAMCH_P / 2.0


def _case(n, seed, frac, gap=0.05):
    """Symmetric A = W diag(lam) W^T, lam in [-1, 1], no eigenvalue within gap of sigma."""
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(-1.0, 1.0, n))
    sigma = float(np.quantile(lam, frac)) if n > 1 else float(lam[0]) + 0.5
    lam = np.where(np.abs(lam - sigma) < gap, lam + np.where(lam >= sigma, gap, -gap), lam)
    A, _ = zo.symmetric_matrix(n, lam, seed + 1)
    return A, sigma


def _check_pairs(A, sigma, w, V, count, tol=1e-12):
    n = A.shape[0]
    ev, W = np.linalg.eigh(A)
    below = ev < sigma
    assert count == int(below.sum())
    assert V.shape == (n, count) and w.shape == (count,)
    if count == 0:
        return
    nrm = np.linalg.norm(A)
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - ev[below])) <= tol * nrm, np.max(np.abs(w - ev[below]))
    res = np.linalg.norm(A @ V - V * w) / nrm
    orth = np.linalg.norm(V.T @ V - np.eye(count))
    assert res <= tol, res
    assert orth <= tol * math.sqrt(n), orth
    Wx = W[:, below]
    sines = np.linalg.svd(V - Wx @ (Wx.T @ V), compute_uv=False)
    assert sines.max() <= 1e-10, sines.max()


@pytest.mark.parametrize("n,frac,precision", [(40, 0.3, "f64"), (100, 0.5, "f64"), (257, 0.2, "f64"),
                                              (300, 0.7, "f64"), (700, 0.45, "f64"), (300, 0.5, "bf16x3"),
                                              (512, 0.35, "bf16x3"), (200, 0.5, "f32")])
def test_eig_below_matches_lapack(zl, torch_gpu, n, frac, precision):
    print(f"eig_below n={n} frac={frac} precision={precision}", flush=True)
    A, sigma = _case(n, 10 + n, frac)
    w, V, count, info, err, iters = zl.eig_below_device(zl.to_device(A), sigma, precision=precision)
    w = w.cpu().numpy()
    V = zl.from_device(V) if count else np.zeros((n, 0))
    # the polar factor in f32 carries ~1e-7 errors into the subspace
    _check_pairs(A, sigma, w, V, count, tol=1e-12 if precision != "f32" else 1e-5)
    assert 1 <= iters <= 2
    if precision != "f32":
        assert err <= 1e-13


@pytest.mark.parametrize("n,frac", [(60, 0.4), (150, 0.6)])
def test_eig_below_matches_oracle(zl, torch_gpu, n, frac):
    print(f"eig_below vs oracle n={n}", flush=True)
    A, sigma = _case(n, 40 + n, frac)
    w_ref, V_ref, log = zo.eig_below(A, sigma, seed=n)
    w, V, count, _, _, _ = zl.eig_below_device(zl.to_device(A), sigma)
    assert count == log.rk
    assert np.allclose(w.cpu().numpy(), w_ref, atol=1e-13 * np.linalg.norm(A))
    V = zl.from_device(V)
    # the same invariant subspace: V V^T = V_ref V_ref^T
    assert np.linalg.norm(V @ V.T - V_ref @ V_ref.T) <= 1e-11


@pytest.mark.parametrize("n", [50, 333])
def test_subit_with_exact_polar_factor(zl, torch_gpu, n):
    """PQDWHSUBIT from the exact sign function: Rk, V orthogonal, E = V^T A V
    block diagonal at the accepted level, A and Q overwritten as the reference."""
    print(f"subit n={n}", flush=True)
    A, sigma = _case(n, 70 + n, 0.5)
    ev, W = np.linalg.eigh(A)
    Q = (W * np.sign(ev - sigma)) @ W.T
    dA, dQ = zl.to_device(A), zl.to_device(Q)
    rk, err, iters = zl.subit_device(dA, float(np.linalg.norm(A)), dQ)
    V, E = zl.from_device(dQ), zl.from_device(dA)
    assert rk == int((ev < sigma).sum())
    assert np.linalg.norm(V.T @ V - np.eye(n)) <= 1e-13 * math.sqrt(n)
    assert np.linalg.norm(E - V.T @ A @ V) <= 1e-13 * np.linalg.norm(A)
    off = np.linalg.norm(E[:rk, rk:]) / np.linalg.norm(A)
    assert off == pytest.approx(err, rel=1e-6, abs=1e-16)
    assert 1 <= iters <= 2 and err <= 1e-13
    if iters == 1:
        assert err <= 4.0 * TOL1


def test_eig_device_from_polar(zl, torch_gpu):
    """ZoloEig with the library's own polar factor of A - sigma I (sym 'S')."""
    n = 220
    A, sigma = _case(n, 91, 0.3)
    B = A - sigma * np.eye(n)
    dB = zl.to_device(B)
    alpha, beta = zl.bounds_device(dB)
    dU, _, _ = zl.pzolopd1_device(dB, alpha, beta, zl.choosem(1.0 / beta), sym="S", polar=False)
    w, V, rk, err, iters = zl.eig_device(dU, zl.to_device(A))
    _check_pairs(A, sigma, w.cpu().numpy(), zl.from_device(V), rk)


@pytest.mark.parametrize("where", ["none", "all"])
def test_eig_below_none_or_all(zl, torch_gpu, where):
    n = 90
    A, _ = _case(n, 5, 0.5)
    ev = np.linalg.eigvalsh(A)
    sigma = float(ev[0] - 0.5) if where == "none" else float(ev[-1] + 0.5)
    w, V, count, _, _, _ = zl.eig_below_device(zl.to_device(A), sigma)
    assert count == (0 if where == "none" else n)
    if count:
        _check_pairs(A, sigma, w.cpu().numpy(), zl.from_device(V), count)


def test_eig_below_order_one(zl, torch_gpu):
    A = np.array([[0.75]])
    w, V, count, _, _, _ = zl.eig_below_device(zl.to_device(A), 1.0)
    assert count == 1 and abs(float(w[0]) - 0.75) < 1e-15
    assert abs(abs(float(zl.from_device(V)[0, 0])) - 1.0) < 1e-15


@pytest.mark.parametrize("bounds", ["qr", "lu"])
def test_eig_below_bounds_modes(zl, torch_gpu, bounds):
    """alpha, beta and r of the polar step of A - sigma I: from the QR its first step
    reuses (oracle bounds_from_qr) or from LU + 1-norm estimate with PQDWHFACS's
    convention (pqdwhfacs.f90:318-344, oracle bounds_estimate(..., 'qdwh'))."""
    n = 160
    A, sigma = _case(n, 91, 0.4)
    B = A - sigma * np.eye(n)
    w, V, count, info, err, iters = zl.eig_below_device(zl.to_device(A), sigma, bounds=bounds)
    alpha, beta = zo.bounds_from_qr(B) if bounds == "qr" else zo.bounds_estimate(B, "qdwh")
    assert abs(info.alpha - alpha) <= 1e-13 * alpha
    assert abs(info.beta - beta) <= 1e-8 * beta, (info.beta, beta)
    assert info.r == zo.choosem(1.0 / beta)
    w = w.cpu().numpy()
    V = zl.from_device(V) if count else np.zeros((n, 0))
    _check_pairs(A, sigma, w, V, count, tol=1e-12)
