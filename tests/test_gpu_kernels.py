"""GPU parity of the building blocks: MFMA GEMM (F64 / F32 / BF16x3), the
recursive Cholesky + inverse, the structured Householder QR, and the single
shift steps ZOLOCHOL / ZOLOQR against the CPU oracle."""
import math

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu


def _gemm_case(torch, zl, prec, ta, tb, m, n, k, beta, seed):
    rng = np.random.default_rng(seed)
    dt = np.float64 if prec in ("f64", "f64x") else np.float32
    Ah = rng.standard_normal((k, m) if ta else (m, k)).astype(dt)
    Bh = rng.standard_normal((n, k) if tb else (k, n)).astype(dt)
    Ch = rng.standard_normal((m, n)).astype(dt)
    dA, dB, dC = zl.to_device(Ah), zl.to_device(Bh), zl.to_device(Ch)
    zl.gemm_device(ta, tb, m, n, k, 1.5, dA, Ah.shape[0], dB, Bh.shape[0], beta, dC, m, precision=prec)
    torch.cuda.synchronize()
    got = zl.from_device(dC)
    opA = Ah.T.astype(np.float64) if ta else Ah.astype(np.float64)
    opB = Bh.T.astype(np.float64) if tb else Bh.astype(np.float64)
    ref = 1.5 * opA @ opB + beta * Ch.astype(np.float64)
    scale = 1.5 * np.abs(opA) @ np.abs(opB) + abs(beta) * np.abs(Ch)
    return float(np.max(np.abs(got - ref) / scale))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("f64x", 1e-14), ("f32", 2e-6), ("bf16x3", 2e-5)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
# BF16X3 products of at most 2^31 M N K batch run on the F32 path; the last
# shape is above that, so the BF16x3 kernel itself sees ragged edge tiles
@pytest.mark.parametrize("m,n,k", [(200, 130, 70), (64, 257, 300), (300, 40, 129), (1, 1, 1), (513, 129, 17),
                                   (256, 256, 33), (1000, 1100, 1037), (1301, 1403, 1207)])
def test_gemm(torch_gpu, zl, prec, tol, ta, tb, m, n, k):
    # tolerance is relative to |A||B| elementwise: a fixed multiple of the unit roundoff
    err = _gemm_case(torch_gpu, zl, prec, ta, tb, m, n, k, 0.5, m * 7 + n + k)
    assert err < tol * max(1, math.sqrt(k)), err


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [37, 64, 200, 513])
def test_chol_inv(torch_gpu, zl, prec, n):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    W = G.T @ G / n + 0.5 * np.eye(n)
    dt = np.float64 if prec in ("f64", "f64x") else np.float32
    dA = zl.to_device(W.astype(dt))
    dLi = zl.to_device(np.zeros((n, n), dtype=dt))
    zl.chol_inv_device(dA, dLi, precision=prec)
    torch_gpu.cuda.synchronize()
    Li = zl.from_device(dLi).astype(np.float64)
    L = np.linalg.cholesky(W)
    tol = 1e-12 if prec == "f64" else 1e-4
    assert np.allclose(np.triu(Li, 1), 0.0)
    assert np.linalg.norm(Li @ L - np.eye(n)) / math.sqrt(n) < tol
    assert np.linalg.norm(Li - np.linalg.inv(L)) / np.linalg.norm(np.linalg.inv(L)) < tol


# n > 544: the panels run as one cooperative launch per 64 columns; n = 1100:
# five 256-wide outer blocks, the last one partial (76 columns)
@pytest.mark.parametrize("n,s", [(5, 1.0), (64, 1e-3), (100, 0.5), (130, 1e4), (257, 2.0), (600, 1e-3), (777, 3.0),
                                 (1100, 0.7)])
def test_stacked_qr(torch_gpu, zl, n, s):
    """MPDGEQRF+MPDORGQR restatement vs the oracle's structured Householder QR."""
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    S = np.zeros((2 * n, n))
    S[:n] = X
    S[n:] = s * np.eye(n)
    dS = zl.to_device(S)
    dQ = zl.to_device(np.zeros((2 * n, n)))
    dtau = torch_gpu.zeros(n, dtype=torch_gpu.float64, device="cuda")
    zl.stacked_qr_device(dS, n, dQ, dtau)
    torch_gpu.cuda.synchronize()
    F = zl.from_device(dS)
    Q = zl.from_device(dQ)
    tau = dtau.cpu().numpy()
    Fo, tauo = zo.structured_geqrf(S, n)
    Qo = zo.structured_orgqr(Fo, tauo, n)
    R = np.triu(F[:n])
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-13 * n
    assert np.linalg.norm(Q @ R - S) / np.linalg.norm(S) < 1e-14 * n
    # same reflectors as the oracle (Householder QR is unique given the sign convention)
    assert np.allclose(tau, tauo, rtol=1e-9, atol=1e-12)
    assert np.linalg.norm(Q - Qo) < 1e-11 * n
    assert np.linalg.norm(np.triu(F[:n]) - np.triu(Fo[:n])) / np.linalg.norm(np.triu(Fo[:n])) < 1e-12 * n


@pytest.mark.parametrize("sym", ["N", "S"])
@pytest.mark.parametrize("n,con,rk", [(96, 1.5, 3), (130, 40.0, 4), (64, 1e3, 2)])
def test_zolochol(zl, sym, n, con, rk):
    rng = np.random.default_rng(n + rk)
    U = zo.haar(n, rng)
    V = U if sym == "S" else zo.haar(n, rng)
    s = np.geomspace(1.0, con, n)
    X = (U * s) @ V.T
    coeff = zo.compute_coeff(con, zo.DLAMCH_P, rk)
    for ip in range(1, rk + 1):
        Zg = zl.zolochol(X, coeff, rk, ip, sym)
        Zo = zo.zolochol_step(X, coeff, rk, ip, sym)
        assert np.linalg.norm(Zg - Zo) / np.linalg.norm(Zo) < 1e-11, ip


@pytest.mark.parametrize("n,con,rk", [(96, 1e6, 5), (130, 1e10, 7), (64, 30.0, 2)])
def test_zoloqr(zl, n, con, rk):
    rng = np.random.default_rng(n + 17)
    U, V = zo.haar(n, rng), zo.haar(n, rng)
    s = np.geomspace(1.0, con, n)
    X = (U * s) @ V.T
    coeff = zo.compute_coeff(con, zo.DLAMCH_P, rk)
    for ip in range(1, rk + 1):
        Zg = zl.zoloqr(X, coeff, rk, ip)
        Zo = zo.zoloqr_step(X, coeff, rk, ip)
        # Z = -a X (X^T X + c I)^{-1} is as sensitive as X^{-T} when c << sigma^2:
        # two backward-stable QRs agree to about eps * kappa(X)
        tol = max(1e-11, 1e-15 * con)
        assert np.linalg.norm(Zg - Zo) / max(np.linalg.norm(Zo), 1e-300) < tol, ip


@pytest.mark.parametrize("k", [1, 3, 13])
def test_gemm_f64x_small_k(torch_gpu, zl, k):
    """int8-modular FP64 products at tiny K, where the inputs are rounded to 59-60 bits
    (scaled integers up to 2^60, the widest the residue split handles): FP64-level error."""
    rng = np.random.default_rng(40 + k)
    m, n = 300, 257
    opA = rng.standard_normal((m, k)) * np.exp2(rng.integers(-30, 30, size=(m, 1)))
    opB = rng.standard_normal((k, n))
    Ch = rng.standard_normal((m, n))
    dA, dB, dC = zl.to_device(opA), zl.to_device(opB), zl.to_device(Ch)
    zl.gemm_device(0, 0, m, n, k, 1.0, dA, m, dB, k, 0.5, dC, m, precision="f64x")
    torch_gpu.cuda.synchronize()
    got = zl.from_device(dC)
    ref = opA @ opB + 0.5 * Ch
    scale = np.abs(opA) @ np.abs(opB) + 0.5 * np.abs(Ch)
    assert float(np.max(np.abs(got - ref) / scale)) < 1e-15


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
def test_gemm_f64x_graded(torch_gpu, zl, ta, tb):
    """int8-modular FP64 products: rows of op(A) and columns of op(B) scaled over 2^+-100,
    zero rows/columns, K not a multiple of the slab; error relative to |A||B| at FP64 level."""
    rng = np.random.default_rng(5 + 2 * ta + tb)
    m, n, k = 390, 271, 1111
    opA = rng.standard_normal((m, k)) * np.exp2(rng.integers(-100, 100, size=(m, 1)))
    opB = rng.standard_normal((k, n)) * np.exp2(rng.integers(-100, 100, size=(1, n)))
    opA[7] = 0.0
    opB[:, 11] = 0.0
    Ah = np.ascontiguousarray(opA.T) if ta else opA
    Bh = np.ascontiguousarray(opB.T) if tb else opB
    Ch = rng.standard_normal((m, n))
    dA, dB, dC = zl.to_device(Ah), zl.to_device(Bh), zl.to_device(Ch)
    zl.gemm_device(ta, tb, m, n, k, -0.75, dA, Ah.shape[0], dB, Bh.shape[0], 0.0, dC, m, precision="f64x")
    torch_gpu.cuda.synchronize()
    got = zl.from_device(dC)
    ref = -0.75 * opA @ opB
    scale = 0.75 * np.abs(opA) @ np.abs(opB)
    assert np.all(got[7] == 0.0) and np.all(got[:, 11] == 0.0)
    mask = scale > 0
    assert float(np.max(np.abs(got - ref)[mask] / scale[mask])) < 1e-14


@pytest.mark.parametrize("ta", [1, 0])
def test_gemm_f64x_one_operand(torch_gpu, zl, ta):
    """int8-modular A^T A and A A^T with one matrix passed as both operands (the residues of
    op(A) are reused for op(B)); the vectors they pair are graded over 2^+-60. Error relative
    to |op(A)||op(B)| at FP64 level, and the result is symmetric."""
    rng = np.random.default_rng(70 + ta)
    m, k = 333, 1201
    X = rng.standard_normal((k, m) if ta else (m, k))
    if ta:
        X *= np.exp2(rng.integers(-60, 60, size=(1, m)))
    else:
        X *= np.exp2(rng.integers(-60, 60, size=(m, 1)))
    opA, opB = (X.T, X) if ta else (X, X.T)
    dX, dC = zl.to_device(X), zl.to_device(np.zeros((m, m)))
    zl.gemm_device(ta, 1 - ta, m, m, k, 1.25, dX, X.shape[0], dX, X.shape[0], 0.0, dC, m, precision="f64x")
    torch_gpu.cuda.synchronize()
    got = zl.from_device(dC)
    ref = 1.25 * opA @ opB
    scale = 1.25 * np.abs(opA) @ np.abs(opB)
    assert float(np.max(np.abs(got - ref) / scale)) < 1e-14
    assert np.array_equal(got, got.T)
