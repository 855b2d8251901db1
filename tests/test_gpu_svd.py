"""Zolo-SVD tail on the GPU: the symmetric eigensolver (ELPA_PDSYEVD2's role,
ZoloCHK1.f90:86) against LAPACK dsyevd (numpy.linalg.eigh), and the SVD
assembly (ZoloCHK1.f90:80-91) against the generator's closed form and the
oracle's zolo_svd."""
import math

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu


def _sym(n, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "gauss":
        B = rng.standard_normal((n, n))
        return 0.5 * (B + B.T)
    Q = zo.haar(n, rng) if n > 1 else np.ones((1, 1))
    if kind == "c4":
        lam = 1.0 + 0.5 * np.sin(np.arange(1, n + 1)) + 1e-3 * rng.standard_normal(n)
    elif kind == "repeated":          # a few distinct values: heavy deflation by rotation
        lam = rng.choice([-1.0, 0.25, 2.0], size=n)
    elif kind == "graded":
        lam = np.geomspace(1.0, 1e-10, n) * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    else:
        raise ValueError(kind)
    return (Q * lam) @ Q.T


@pytest.mark.parametrize("n,kind", [(1, "gauss"), (2, "gauss"), (3, "c4"), (17, "gauss"), (32, "c4"), (33, "repeated"),
                                    (64, "graded"), (100, "gauss"), (257, "c4"), (300, "repeated"), (600, "c4"),
                                    (1537, "graded"), (2000, "gauss")])
def test_syevd_matches_lapack(zl, torch_gpu, n, kind):
    H = _sym(n, kind, 3 + n)
    dH = zl.to_device(H)
    w, Z, _ = zl.syevd_device(dH)
    w = w.cpu().numpy()
    Z = zl.from_device(Z)
    wref = np.linalg.eigvalsh(H)
    nrm = max(np.abs(wref).max(), 1e-300)
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - wref)) <= 2e-13 * math.sqrt(n) * nrm, np.max(np.abs(w - wref))
    res = np.linalg.norm(H @ Z - Z * w) / (np.linalg.norm(H) + 1e-300)
    orth = np.linalg.norm(Z.T @ Z - np.eye(n)) / math.sqrt(n)
    assert res <= 2e-13 * math.sqrt(n), res
    assert orth <= 2e-13 * math.sqrt(n), orth


def test_syevd_diagonal_and_identity(zl, torch_gpu):
    for H in (np.diag(np.arange(40, 0, -1, dtype=np.float64)), np.eye(50)):
        w, Z, _ = zl.syevd_device(zl.to_device(H))
        w = w.cpu().numpy()
        Z = zl.from_device(Z)
        assert np.allclose(w, np.sort(np.diag(H)), atol=1e-13)
        assert np.linalg.norm(H @ Z - Z * w) < 1e-12
        assert np.linalg.norm(Z.T @ Z - np.eye(H.shape[0])) < 1e-12


@pytest.mark.parametrize("n", [96, 256])
def test_zolo_svd_c4_closed_form(zl, torch_gpu, n):
    rng = np.random.default_rng(40 + n)
    U0 = zo.haar(n, rng)
    V0 = zo.haar(n, rng)
    s = 1.0 + 0.5 * np.sin(np.arange(1, n + 1)) + 1e-3 * rng.standard_normal(n)
    A = (U0 * s) @ V0.T
    S, dU, dV, info, _ = zl.svd_device(zl.to_device(A))
    S = S.cpu().numpy()
    U = zl.from_device(dU)
    V = zl.from_device(dV)
    assert np.max(np.abs(S - np.sort(s))) <= 1e-13
    assert np.linalg.norm(A - (U * S) @ V.T) / np.linalg.norm(A) <= 1e-13
    # the reference's own residual (check_routines.f90:70-86), within 10x of
    # the oracle's on the same matrix
    Sref, Uref, Vref = zo.zolo_svd(A)
    res, res_o = zo.svd_residual(A, U, S, V), zo.svd_residual(A, Uref, Sref, Vref)
    assert res <= 10 * max(res_o, 1e-15), (res, res_o)
    assert np.linalg.norm(U.T @ U - np.eye(n)) / math.sqrt(n) <= 1e-13
    assert np.linalg.norm(V.T @ V - np.eye(n)) / math.sqrt(n) <= 1e-13
    assert np.max(np.abs(S - Sref)) <= 1e-13


@pytest.mark.parametrize("bounds", ["qr", "lu"])
def test_zolo_svd_bounds_modes(zl, torch_gpu, bounds):
    """alpha, beta and r of the SVD's polar step: from the QR its first step reuses
    (oracle bounds_from_qr) or from LU + 1-norm estimate as the reference's driver
    computes them (test_zolodrv1.f90:179-189, oracle bounds_estimate); r = choosem(1/beta)."""
    n = 200
    rng = np.random.default_rng(77)
    U0, V0 = zo.haar(n, rng), zo.haar(n, rng)
    s = np.geomspace(1.0, 1e-6, n)
    A = (U0 * s) @ V0.T
    S, dU, dV, info, _ = zl.svd_device(zl.to_device(A), bounds=bounds)
    alpha, beta = zo.bounds_from_qr(A) if bounds == "qr" else zo.bounds_estimate(A)
    assert abs(info.alpha - alpha) <= 1e-13 * alpha
    assert abs(info.beta - beta) <= 1e-8 * beta, (info.beta, beta)
    assert info.r == zo.choosem(1.0 / beta)
    S = S.cpu().numpy()
    U, V = zl.from_device(dU), zl.from_device(dV)
    assert np.max(np.abs(S - np.sort(s))) <= 1e-13
    assert np.linalg.norm(A - (U * S) @ V.T) / np.linalg.norm(A) <= 1e-13
