"""Zolo-PD with the bounds taken from the QR of A/alpha (zolo_pd_auto_device):
alpha and beta against the oracle's bounds_from_qr (LAPACK-pinned), the polar
factor against the generator's closed form U V^T and against zolo_pd_device
run with the same (alpha, beta), in FP64 and the mixed precisions."""
import math

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu


def _run(zl, A, precision="f64", rk=0, sym="N"):
    dU, dH, info = zl.pzolopd_auto_device(zl.to_device(A), precision=precision, rk=rk, sym=sym)
    return zl.from_device(dU), zl.from_device(dH), info


@pytest.mark.parametrize("n,kind,kappa", [(1, "geometric", 1.0), (2, "geometric", 10.0), (17, "geometric", 1e4),
                                          (100, "geometric", 1e6), (300, "cluster", 1e8), (700, "geometric", 1e12),
                                          (1500, "cluster", 1e10)])
def test_auto_bounds_match_oracle_and_closed_form(zl, torch_gpu, n, kind, kappa):
    print(f"auto bounds n={n} {kind} kappa={kappa:g}", flush=True)
    A, U0, s, V0 = zo.make_matrix(n, kind, kappa, 300 + n)
    U, H, info = _run(zl, A)
    ao, bo = zo.bounds_from_qr(A)
    assert info.alpha == pytest.approx(ao, rel=1e-13)
    assert info.beta == pytest.approx(bo, rel=1e-8)
    assert info.r == zl.choosem(1.0 / info.beta)
    assert info.beta <= s.min() / ao * (1 + 1e-10)
    # forming A in FP64 already moves its polar factor by about eps * kappa
    Up = U0 @ V0.T
    assert np.linalg.norm(U - Up) / math.sqrt(n) <= max(1e-13, 1e-14 * kappa)
    assert zo.orthogonality(U)[0] <= 1e-14 * max(1, math.sqrt(n) / 8)
    assert zo.backward_error(A, U, H) <= 1e-14


@pytest.mark.parametrize("n,precision", [(700, "f64"), (700, "bf16x3"), (1100, "f32")])
def test_auto_equals_explicit_bounds(zl, torch_gpu, n, precision):
    """The same iteration as zolo_pd_device given the auto (alpha, beta): only
    the rounding of the R scaling differs."""
    print(f"auto vs explicit n={n} {precision}", flush=True)
    A, *_ = zo.make_matrix(n, "geometric", 1e8, 900 + n)
    U, H, info = _run(zl, A, precision)
    dA = zl.to_device(A)
    dU2, dH2, info2 = zl.pzolopd1_device(dA, info.alpha, info.beta, info.r, precision=precision)
    U2 = zl.from_device(dU2)
    assert info.schedule == info2.schedule
    # kappa = 1e8: the polar factor moves by ~eps * kappa under any change of rounding
    assert np.linalg.norm(U - U2) / math.sqrt(n) <= 1e-16 * 1e8
    assert zo.orthogonality(U)[0] <= 1e-14 and zo.backward_error(A, U, H) <= 1e-14


def test_auto_given_rk_and_symmetric(zl, torch_gpu):
    n = 257
    rng = np.random.default_rng(7)
    W = zo.haar(n, rng)
    lam = np.geomspace(1.0, 1e-6, n) * np.where(np.arange(n) % 3 == 0, -1.0, 1.0)
    A = (W * lam) @ W.T
    A = 0.5 * (A + A.T)
    U, H, info = _run(zl, A, rk=3, sym="S")
    assert info.r == 3
    Q = (W * np.sign(lam)) @ W.T
    assert np.linalg.norm(U - Q) / math.sqrt(n) <= 1e-14 * 1e6   # eps * kappa, kappa = 1e6


@pytest.mark.parametrize("n,zc", [(300, 7), (700, 650)])
def test_auto_exactly_singular(zl, torch_gpu, n, zc):
    """A zero column gives a zero diagonal entry of R: no positive bound, an error."""
    A, *_ = zo.make_matrix(n, "geometric", 1e3, n + zc)
    A[:, zc] = 0.0
    with pytest.raises(zl.ZoloError, match="numerically singular"):
        _run(zl, A)
