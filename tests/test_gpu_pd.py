"""GPU parity of the polar-decomposition drivers against the CPU oracle and
the closed form of the generator (A = U diag(s) V^T  =>  U_p = U V^T,
H = V diag(s) V^T)."""
import math

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu


def _errs(A, U, H):
    return zo.backward_error(A, U, H), zo.orthogonality(U)[0]


# n > 544 takes the factored QR step (X = Q_X R once, then [R; s I] per shift)
@pytest.mark.parametrize("n,kind,kappa,seed", [(64, "geometric", 1e6, 1), (200, "loguniform", 1e12, 2),
                                               (150, "cluster", 1e8, 3), (128, "sine", 0, 4),
                                               (700, "loguniform", 1e12, 11), (600, "cluster", 1e8, 12)])
def test_pzolopd1_f64_matches_oracle(zl, n, kind, kappa, seed):
    A, Ug, s, Vg = zo.make_matrix(n, kind, kappa, seed)
    alpha, beta = zo.bounds_estimate(A)
    rk = zo.choosem(1.0 / beta)
    Uo, lo = zo.pzolopd1(A, alpha, beta, rk)
    U, H, info = zl.pzolopd1(A, alpha, beta, rk)
    # same schedule and iteration count as the reference restatement
    sched = "".join("Q" if x == "QR" else "C" for x in lo.schedule)
    assert info.schedule == sched and info.iters == lo.iters
    assert info.shifts == [rk] * info.iters
    assert np.allclose(info.con, lo.con, rtol=1e-9)
    # FormX2's ERR = ||X+ - X||_F of every iteration (FormX2.f90:85), absolute
    # as in the reference.  After a QR step at CON ~ 1e14 both iterates carry
    # forward errors near 1e-8 relative (measured 4e-8 on the GPU at n = 700),
    # and the last ERR is rounding noise; a dropped 1/FONE or ERR^2 for ERR is
    # off by orders of magnitude more
    assert np.allclose(info.err, lo.err, rtol=1e-6, atol=1e-12 * math.sqrt(n)), (info.err, lo.err)
    Ho = zo.polar_h(Uo, A)
    be, orth = _errs(A, U, H)
    beo, ortho = _errs(A, Uo, Ho)
    assert be <= 10 * max(beo, 1e-15) and orth <= 10 * max(ortho, 1e-15), (be, beo, orth, ortho)
    # distance to the oracle's polar factor and to the closed form
    d_o = np.linalg.norm(U - Uo) / math.sqrt(n)
    d_cf = np.linalg.norm(U - Ug @ Vg.T) / math.sqrt(n)
    d_cf_o = np.linalg.norm(Uo - Ug @ Vg.T) / math.sqrt(n)
    assert d_cf <= 10 * max(d_cf_o, 1e-14), (d_cf, d_cf_o)
    assert d_o <= 10 * max(d_cf_o, 1e-14)
    assert np.allclose(H, H.T)


# orders below and just past the 64-wide blocks, one to eight shifts
@pytest.mark.parametrize("n,rk", [(1, 1), (2, 2), (3, 8), (17, 3), (65, 5), (129, 1)])
def test_pzolopd1_small_and_ragged(zl, n, rk):
    A, Ug, s, Vg = zo.make_matrix(n, "geometric", 1e4, 100 + n)
    alpha, beta = zo.bounds_estimate(A)
    Uo, lo = zo.pzolopd1(A, alpha, beta, rk)
    U, H, info = zl.pzolopd1(A, alpha, beta, rk)
    sched = "".join("Q" if x == "QR" else "C" for x in lo.schedule)
    assert info.schedule == sched and info.iters == lo.iters
    be, orth = _errs(A, U, H)
    beo, ortho = _errs(A, Uo, zo.polar_h(Uo, A))
    assert be <= 10 * max(beo, 1e-15) and orth <= 10 * max(ortho, 1e-15), (be, beo, orth, ortho)
    d_cf_o = np.linalg.norm(Uo - Ug @ Vg.T) / math.sqrt(n)
    assert np.linalg.norm(U - Ug @ Vg.T) / math.sqrt(n) <= 10 * max(d_cf_o, 1e-14)


# SYM = 'S' (zolochol.f90:99) on a symmetric indefinite A = Q diag(l) Q^T,
# whose polar factor is Q sign(l) Q^T
@pytest.mark.parametrize("n,rk", [(17, 3), (65, 5), (200, 4)])
def test_pzolopd1_symmetric(zl, n, rk):
    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1, 1e-4, n) * np.where(np.arange(n) % 3 == 0, -1.0, 1.0)
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    alpha, beta = zo.bounds_estimate(A)
    Uo, lo = zo.pzolopd1(A, alpha, beta, rk, sym="S")
    U, H, info = zl.pzolopd1(A, alpha, beta, rk, sym="S")
    sched = "".join("Q" if x == "QR" else "C" for x in lo.schedule)
    assert info.schedule == sched and info.iters == lo.iters
    Ucf = (Q * np.sign(lam)) @ Q.T
    d_o = np.linalg.norm(Uo - Ucf) / math.sqrt(n)
    assert np.linalg.norm(U - Ucf) / math.sqrt(n) <= 10 * max(d_o, 1e-14), d_o


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("n,kind,kappa,seed", [(256, "cluster", 1e8, 5), (200, "loguniform", 1e12, 6),
                                               (650, "cluster", 1e8, 13)])
def test_pzolopd1_mixed(zl, prec, n, kind, kappa, seed):
    A, Ug, s, Vg = zo.make_matrix(n, kind, kappa, seed)
    alpha, beta = zo.bounds_estimate(A)
    rk = zo.choosem(1.0 / beta)
    U, H, info = zl.pzolopd1(A, alpha, beta, rk, precision=prec)
    assert info.schedule[0] == "Q" and set(info.schedule[1:]) <= {"c"}
    assert info.polish_passes >= 1
    # the QR step keeps r; a low-precision step takes the fewest shifts whose
    # bound reaches FP32 convergence in one step (scalar recurrence of the oracle)
    assert info.shifts[0] == rk
    for k in range(1, info.iters):
        rs = info.shifts[k]
        assert 1 <= rs <= rk
        _, c1 = zo.zolo_fone_con(info.con[k], zo.compute_coeff(info.con[k], zo.DLAMCH_P, rs), rs)
        assert c1 - 1.0 <= 6e-8 or rs == rk
        if rs > 1:
            _, c0 = zo.zolo_fone_con(info.con[k], zo.compute_coeff(info.con[k], zo.DLAMCH_P, rs - 1), rs - 1)
            assert c0 - 1.0 > 6e-8
    be, orth = _errs(A, U, H)
    assert be <= 1e-13 and orth <= 1e-13, (be, orth)
    Uo, _ = zo.pzolopd1(A, alpha, beta, rk)
    d_cf = np.linalg.norm(U - Ug @ Vg.T) / math.sqrt(n)
    d_cf_o = np.linalg.norm(Uo - Ug @ Vg.T) / math.sqrt(n)
    assert d_cf <= 10 * max(d_cf_o, 1e-14), (d_cf, d_cf_o)


@pytest.mark.parametrize("n,kappa,seed", [(64, 1e6, 7), (160, 1e10, 8), (100, 10.0, 9)])
def test_pqdwhfacs_matches_oracle(zl, n, kappa, seed):
    A, Ug, s, Vg = zo.make_matrix(n, "geometric", kappa, seed)
    Uo, Ho, lo = zo.pqdwhfacs(A)
    U, H, info = zl.pqdwhfacs(A)
    sched = "".join("Q" if x == "QR" else "C" for x in lo.schedule)
    assert info.schedule == sched
    assert abs(info.alpha - lo.alpha) <= 1e-13 * lo.alpha
    # the 1-norm estimate runs on a different LU (blocked GPU vs LAPACK): it
    # agrees to about kappa * eps
    assert abs(info.beta - lo.beta) <= max(1e-8, 1e-15 * kappa) * lo.beta
    be, orth = _errs(A, U, H)
    beo, ortho = _errs(A, Uo, Ho)
    assert be <= 10 * max(beo, 1e-15) and orth <= 10 * max(ortho, 1e-15)
    assert np.linalg.norm(U - Uo) <= 10 * max(np.linalg.norm(Uo - Ug @ Vg.T), 1e-13)


# n > 512: the LU panels span several workgroups (one cooperative launch each)
@pytest.mark.parametrize("n,conv", [(50, "zolo"), (300, "zolo"), (257, "qdwh"), (1100, "zolo"), (1500, "qdwh")])
def test_bounds(torch_gpu, zl, n, conv):
    A, *_ = zo.make_matrix(n, "geometric", 1e7, n)
    dA = zl.to_device(A)
    a, b = zl.bounds_device(dA, conv)
    ao, bo = zo.bounds_estimate(A, conv)
    assert abs(a - ao) <= 1e-13 * ao
    assert abs(b - bo) <= max(1e-8, 1e-15 * 1e7) * bo, (b, bo)


@pytest.mark.parametrize("n,zc", [(300, 7), (700, 70), (700, 650)])
def test_bounds_exactly_singular(torch_gpu, zl, n, zc):
    """A zero column stays exactly zero through the elimination, so its pivot is
    0 (dgetrf INFO = column + 1) and the bound estimate refuses the matrix.
    n = 700: the LU panels of the first columns span two workgroups."""
    A, *_ = zo.make_matrix(n, "geometric", 1e3, n + zc)
    A[:, zc] = 0.0
    with pytest.raises(zl.ZoloError, match=rf"exactly singular \(LU pivot {zc + 1}\)"):
        zl.bounds_device(zl.to_device(A), "zolo")


@pytest.mark.parametrize("m,n", [(100, 100), (300, 120), (77, 200)])
def test_normest(torch_gpu, zl, m, n):
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    dA = zl.to_device(A)
    est, cnt = zl.pdnormest_device(dA, 1e-8)
    eo, co = zo.normest(A, 1e-8)
    assert abs(est - eo) <= 1e-10 * eo and cnt == co
    # power iteration approaches ||A||_2 from below
    assert est <= np.linalg.norm(A, 2) * (1 + 1e-12) and est >= 0.99 * np.linalg.norm(A, 2)


def test_pzolopd1_device_entry(torch_gpu, zl):
    n = 96
    A, Ug, s, Vg = zo.make_matrix(n, "geometric", 1e5, 11)
    alpha, beta = zo.bounds_estimate(A)
    dA = zl.to_device(A)
    dU, dH, info = zl.pzolopd1_device(dA, alpha, beta, 5)
    torch_gpu.cuda.synchronize()
    U = zl.from_device(dU)
    Uh, _, _ = zl.pzolopd1(A, alpha, beta, 5)
    assert np.linalg.norm(U - Uh) < 1e-13 * n
    assert info.time_ms > 0


def test_argument_errors(zl):
    A = np.eye(4)
    with pytest.raises(zl.ZoloError) as e:
        zl.pzolopd1(A, 1.0, 0.5, 9)
    assert e.value.status == -3
    with pytest.raises(zl.ZoloError) as e:
        zl.pzolopd1(A, 1.0, 2.0, 2)
    assert e.value.status == -6


def test_pzolopd2_restart_matches_oracle(zl):
    """PZOLOPD2 restart (pzolopd2.f90:397-405): CON collapses to 1 after the first
    step of a numerically singular matrix; both the GPU and the oracle restart
    from the iterate with re-estimated bounds and reach the SVD polar factor."""
    rng = np.random.default_rng(23)
    n = 120
    A = (zo.haar(n, rng) * np.geomspace(1.0, 1e-3, n)) @ zo.haar(n, rng).T
    A[:, 1] = A[:, 0]
    A[7, 0] = 0.0
    A[7, 1] = 2e-14
    al, be = zo.bounds_estimate(A)
    rk = zo.choosem(1.0 / be)
    Uo, lo = zo.pzolopd2(A, al, be, rk)
    U, H, info = zl.pzolopd2(A, al, be, rk)
    assert lo.restarts >= 1 and info.restarts == lo.restarts
    assert info.schedule == "".join("Q" if x == "QR" else "C" for x in lo.schedule)
    W, s, Vt = np.linalg.svd(A)
    assert np.linalg.norm(U - W @ Vt) / np.sqrt(n) < 1e-12
    assert zo.orthogonality(U)[0] < 1e-13


# FormX2 (FormX2.f90:73-89) through its own C-ABI entry point: X+ = (X + Z) *
# (1/FONE) element by element, as PDLASCL scales (bit for bit with the host),
# and ERR = ||X+ - X||_F
@pytest.mark.parametrize("n,fone", [(1, 1.0), (17, 1.75), (300, 0.6), (1025, 3.0e3)])
def test_formx2_matches_oracle(zl, n, fone):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    Z = rng.standard_normal((n, n)) * 1e-3
    Xo, eo = zo.formx2(X, Z, fone)
    Xg, eg = zl.formx2(X, Z, fone)
    assert np.array_equal(Xg, Xo)
    assert math.isclose(eg, eo, rel_tol=1e-13), (eg, eo)
