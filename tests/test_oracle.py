"""Pins of the CPU oracle (oracle/zolo_oracle.py) against mathematics that
fixes each function independently of its own formula: mpmath's elliptic
functions, the equioscillation and partial-fraction identities of the
Zolotarev rational function, the optimality property of the QDWH weights,
LAPACK (QR, LU condition estimate) and scipy's polar decomposition, and the
closed form of the generated test matrices.  Mutation tests plant plausible
mistakes in copies of oracle functions and check that a pin catches them."""
import math

import mpmath as mp
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import zolo_oracle as zo

EPS = zo.DLAMCH_P


# ---------------------------------------------------------------- elliptic functions
@pytest.mark.parametrize("alpha", [0.05, 0.7, 1.3, math.acos(1e-4), math.acos(1e-9)])
def test_mellipke_vs_mpmath(alpha):
    mp.mp.dps = 40
    m = mp.sin(alpha) ** 2
    k, e = zo.mellipke(alpha, EPS)
    assert k == pytest.approx(float(mp.ellipk(m)), rel=1e-14)
    assert e == pytest.approx(float(mp.ellipe(m)), rel=1e-10, abs=1e-13)


@pytest.mark.parametrize("alpha", [0.3, 1.2, math.acos(1e-6)])
@pytest.mark.parametrize("frac", [0.1, 0.45, 0.93])
def test_mellipj_vs_mpmath(alpha, frac):
    mp.mp.dps = 40
    m = mp.sin(alpha) ** 2
    u = frac * float(mp.ellipk(m))
    sn, cn, dn = zo.mellipj(u, alpha, EPS)
    assert sn == pytest.approx(float(mp.ellipfun("sn", u, m=m)), rel=1e-12, abs=1e-15)
    # cn and dn lose relative accuracy when they are small (m -> 1, u -> K): the
    # descending Landen recurrence of Auxil_Zolo.f90:116-127 carries an absolute error
    assert cn == pytest.approx(float(mp.ellipfun("cn", u, m=m)), rel=1e-10, abs=1e-10)
    assert dn == pytest.approx(float(mp.ellipfun("dn", u, m=m)), rel=1e-8, abs=1e-10)


# ---------------------------------------------------------------- Zolotarev coefficients
def _zolo_rational(x, coeff, rk):
    x = np.asarray(x, dtype=float)
    f = x.copy()
    for j in range(rk):
        f *= (x * x + coeff[2 * j + 1]) / (x * x + coeff[2 * j])
    return f


@pytest.mark.parametrize("con,rk", [(3.0, 2), (150.0, 3), (1e4, 4), (1e6, 5), (1e8, 6), (1e12, 7), (1.4, 6)])
def test_coeff_vs_mpmath(con, rk):
    """c_i = sn^2/cn^2(i K/(2r+1); m) at 40 digits.

    The reference parametrises the modulus by alpha = acos(1/CON) in double
    precision (Auxil_Zolo.f90:176-177), so the modulus it actually uses is
    m = sin(alpha)^2 of that double alpha; for CON = 1e12 this differs from
    1 - 1/CON^2 by ~1e-4 in k'.  The pin uses the same alpha.
    """
    mp.mp.dps = 40
    alpha = math.acos(1.0 / con)
    m = mp.sin(mp.mpf(alpha)) ** 2
    K = mp.ellipk(m)
    coeff = zo.compute_coeff(con, EPS, rk)
    for i in range(1, 2 * rk + 1):
        u = i * K / (2 * rk + 1)
        sn, cn = mp.ellipfun("sn", u, m=m), mp.ellipfun("cn", u, m=m)
        # the Landen recurrence gives cn to an absolute accuracy, so c_i = sn^2/cn^2
        # is relatively accurate to ~1e-11/cn (measured up to 6e-12 absolute in cn)
        tol = max(1e-8, 3e-11 / abs(float(cn)))
        assert coeff[i - 1] == pytest.approx(float(sn ** 2 / cn ** 2), rel=tol)


@pytest.mark.parametrize("con,rk", [(3.0, 2), (150.0, 3), (1e6, 5), (1e8, 6), (1e12, 7)])
def test_partial_fraction_identity(con, rk):
    """prod (x^2 + c_2j)/(x^2 + c_2j-1) = 1 - sum a_j/(x^2 + c_2j-1): pins a_j and its sign."""
    coeff = zo.compute_coeff(con, EPS, rk)
    x = np.geomspace(1.0, con, 37)
    prod = _zolo_rational(x, coeff, rk) / x
    pf = 1.0 - sum(coeff[2 * rk + j] / (x * x + coeff[2 * j]) for j in range(rk))
    assert np.allclose(pf, prod, rtol=1e-9)


@pytest.mark.parametrize("con,rk", [(3.0, 2), (40.0, 3), (1e4, 4), (1e8, 6)])
def test_equioscillation_and_con_update(con, rk):
    """f/FONE maps [1, CON] into [1, CON_next], touching both ends (best rational approximation)."""
    coeff = zo.compute_coeff(con, EPS, rk)
    fone, con_next = zo.zolo_fone_con(con, coeff, rk)
    x = np.geomspace(1.0, con, 200001)
    g = _zolo_rational(x, coeff, rk) / fone
    # the double-precision alpha = acos(1/CON) perturbs the modulus by ~eps*CON
    slack = max(1e-10, 1e-15 * con)
    assert g.min() >= 1.0 - slack
    assert g.max() <= con_next * (1 + slack)
    assert g.max() >= con_next * (1 - 1e-6)
    # number of interior local maxima reaching CON_next: r of them (equioscillation)
    interior = (g[1:-1] > g[:-2]) & (g[1:-1] > g[2:]) & (g[1:-1] > con_next * (1 - 1e-4))
    assert interior.sum() == rk


def test_choosem_table():
    cases = {1.0005: 2, 1.005: 3, 1.05: 4, 1.15: 5, 1.4: 6, 1.9: 8, 5.0: 2, 100.0: 3, 1e4: 4, 1e6: 5,
             1e8: 6, 1e12: 7, 1e13: 8}
    for con, r in cases.items():
        assert zo.choosem(con) == r


# ---------------------------------------------------------------- QDWH weights
def _qdwh_min(a, ell, x):
    b = (a - 1) ** 2 / 4
    c = a + b - 1
    g = x * (a + b * x * x) / (1 + c * x * x)
    return g, b, c


@pytest.mark.parametrize("ell", [1e-9, 1e-6, 1e-3, 0.1, 0.7, 0.99])
def test_hfunc_optimality(ell):
    """a = h(l) maximises min_{[l,1]} x(a+bx^2)/(1+cx^2) while keeping it <= 1 (NBG 2010)."""
    x = np.geomspace(ell, 1.0, 400001)
    a = zo.hfunc(ell)
    g, b, c = _qdwh_min(a, ell, x)
    assert g.max() <= 1 + 1e-9
    # the minimum over [l, 1] is attained at l: it is the next lower bound (pqdwhfacs.f90:366-367)
    assert g.min() == pytest.approx(ell * (a + b * ell * ell) / (1 + c * ell * ell), rel=1e-6)
    for f in (0.97, 1.03):
        g2, _, _ = _qdwh_min(a * f, ell, x)
        assert g2.min() < g.min() * (1 - 1e-9) or g2.max() > 1 + 1e-9


def test_hfunc_fixed_point():
    assert zo.hfunc(1.0) == pytest.approx(3.0, rel=1e-15)


# ---------------------------------------------------------------- structured QR
@pytest.mark.parametrize("n,s", [(1, 2.0), (7, 1e-3), (33, 1.0), (40, 1e5)])
def test_structured_qr_vs_lapack(n, s):
    rng = np.random.default_rng(n)
    S = np.zeros((2 * n, n))
    S[:n] = rng.standard_normal((n, n))
    S[n:] = s * np.eye(n)
    F, tau = zo.structured_geqrf(S, n)
    Q = zo.structured_orgqr(F, tau, n)
    R = np.triu(F[:n])
    Ql, Rl = np.linalg.qr(S)
    d = np.sign(np.diag(R)) * np.sign(np.diag(Rl))
    assert np.allclose(R, d[:, None] * Rl, rtol=1e-10, atol=1e-12 * np.abs(Rl).max())
    assert np.allclose(Q, Ql * d[None, :], atol=1e-12)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-13 * max(n, 1)
    assert np.linalg.norm(Q @ R - S) < 1e-13 * np.linalg.norm(S) * max(n, 1)
    # structure: Q2 = s R^{-1} is upper triangular
    assert np.allclose(np.tril(Q[n:], -1), 0, atol=1e-14)


# ---------------------------------------------------------------- shift steps
@pytest.mark.parametrize("con,rk", [(30.0, 3), (1e5, 5)])
def test_shift_steps_closed_form(con, rk):
    """Both shift formulas against Z = -a X (X^T X + c I)^{-1} evaluated at 40 digits."""
    n = 16
    rng = np.random.default_rng(3)
    X = zo.haar(n, rng) @ np.diag(np.geomspace(1, con, n)) @ zo.haar(n, rng).T
    coeff = zo.compute_coeff(con, EPS, rk)
    mp.mp.dps = 40
    Xm = mp.matrix(X.tolist())
    G = Xm.T * Xm
    for ip in range(1, rk + 1):
        c = coeff[2 * ip - 2]
        Wm = G + mp.mpf(c) * mp.eye(n)
        Zm = -mp.mpf(coeff[2 * rk + ip - 1]) * (Xm * mp.inverse(Wm))
        ref = np.array(Zm.tolist(), dtype=float)
        zc = zo.zolochol_step(X, coeff, rk, ip)
        zq = zo.zoloqr_step(X, coeff, rk, ip)
        cond = (con * con + c) / (1 + c)   # condition of X^T X + c I (Cholesky route)
        assert np.linalg.norm(zc - ref) <= max(1e-12, 1e-15 * cond) * np.linalg.norm(ref)
        # the QR route is sensitive only to kappa(X), not its square
        assert np.linalg.norm(zq - ref) <= max(1e-12, 1e-14 * con) * np.linalg.norm(ref)


def test_lapack_qr_step_matches_structured():
    """The LAPACK-QR form of the QR step (the timed CPU baseline) gives the
    structured step's Z, and the driver built on it the same schedule and
    polar factor."""
    n, rk = 24, 4
    rng = np.random.default_rng(11)
    X = zo.haar(n, rng) @ np.diag(np.geomspace(1, 1e5, n)) @ zo.haar(n, rng).T
    coeff = zo.compute_coeff(1e5, EPS, rk)
    for ip in range(1, rk + 1):
        a = zo.zoloqr_step(X, coeff, rk, ip)
        b = zo.zoloqr_step_lapack(X, coeff, rk, ip)
        assert np.linalg.norm(a - b) <= 1e-14 * 1e5 * np.linalg.norm(a)   # the QR route: u * kappa(X)
    A, U, s, V = zo.make_matrix(40, "loguniform", 1e12, 9)
    al, be = zo.bounds_estimate(A)
    rk = zo.choosem(1 / be)
    Q1, l1 = zo.pzolopd1(A, al, be, rk)
    Q2, l2 = zo.pzolopd1(A, al, be, rk, lapack_qr=True)
    assert l1.schedule == l2.schedule and l1.iters == l2.iters
    assert np.linalg.norm(Q1 - Q2) < 1e-13 * 1e12 * 40 / 1e4


# ---------------------------------------------------------------- drivers vs closed form
@pytest.mark.parametrize("n,kind,kappa", [(40, "geometric", 1e6), (48, "loguniform", 1e12), (40, "cluster", 1e8),
                                          (32, "sine", 0)])
def test_pzolopd1_closed_form(n, kind, kappa):
    A, U, s, V = zo.make_matrix(n, kind, kappa, 5)
    al, be = zo.bounds_estimate(A)
    rk = zo.choosem(1 / be)
    Q, log = zo.pzolopd1(A, al, be, rk)
    H = zo.polar_h(Q, A)
    Up, Hp = sla.polar(A)
    assert zo.backward_error(A, Q, H) < 1e-14
    assert zo.orthogonality(Q)[0] < 1e-14
    cond = s.max() / s.min()
    assert np.linalg.norm(Q - U @ V.T) < 1e-14 * cond * n
    assert np.linalg.norm(Q - Up) < 1e-14 * cond * n
    assert log.schedule[0] == ("QR" if kind != "sine" else log.schedule[0])


def test_pzolopd_schedule_kappa_1e12():
    """kappa=1e12: r=7, QR then Cholesky steps, three iterations (absolute ERR test)."""
    A, U, s, V = zo.make_matrix(64, "loguniform", 1e12, 2)
    al, be = zo.bounds_estimate(A)
    rk = zo.choosem(1 / be)
    _, log = zo.pzolopd1(A, al, be, rk)
    assert rk in (7, 8)
    assert log.schedule == ["QR", "CHOL", "CHOL"]


@pytest.mark.parametrize("n,kappa", [(40, 1e6), (30, 1e10), (32, 5.0)])
def test_pqdwhfacs_closed_form(n, kappa):
    A, U, s, V = zo.make_matrix(n, "geometric", kappa, 9)
    Q, H, log = zo.pqdwhfacs(A)
    Up, Hp = sla.polar(A)
    assert zo.backward_error(A, Q, H) < 1e-14
    assert np.linalg.norm(Q - Up) < 1e-14 * kappa * n
    assert np.allclose(H, Hp, atol=1e-13 * np.abs(Hp).max())
    assert len(log.schedule) <= 7


# ---------------------------------------------------------------- bounds, normest
@pytest.mark.parametrize("n", [2, 9, 60])
def test_lacn2_matches_dgecon(n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    lp = sla.lu_factor(X)
    est, _ = zo.lacn2_norm1_inv(lp)
    rc, info = sla.lapack.dgecon(lp[0], 1.0, norm="1")
    assert est == pytest.approx(1.0 / rc, rel=1e-12)
    assert est <= np.abs(np.linalg.inv(X)).sum(axis=0).max() * (1 + 1e-12)


@pytest.mark.parametrize("n", [2, 9, 60, 200])
def test_lacn2_r_matches_dtrcon(n):
    rng = np.random.default_rng(100 + n)
    R = np.linalg.qr(rng.standard_normal((n, n)), mode="r")
    est, _ = zo.lacn2_norm1_inv_r(R)
    rc, info = sla.lapack.dtrcon(R, norm="1", uplo="U", diag="N")
    assert info == 0
    assert est == pytest.approx(1.0 / (rc * np.abs(R).sum(axis=0).max()), rel=1e-10)
    assert est <= np.abs(np.linalg.inv(R)).sum(axis=0).max() * (1 + 1e-12)


@pytest.mark.parametrize("kind,kappa", [("geometric", 1e7), ("cluster", 1e10), ("sine", 1.0)])
def test_beta_from_qr_is_lower_bound(kind, kappa):
    A, U, s, V = zo.make_matrix(80, kind, kappa, 5)
    al, be = zo.bounds_from_qr(A)
    assert al == pytest.approx(np.linalg.norm(A))
    smin = s.min() / al
    assert be <= smin * (1 + 1e-12)
    assert be >= smin / (10 * math.sqrt(80))
    # the estimate on R and the one on the LU factors agree to within the 1-norm's sqrt(n) slack
    _, be_lu = zo.bounds_estimate(A)
    assert be_lu / math.sqrt(80) <= be <= be_lu * math.sqrt(80)


def test_beta_is_lower_bound():
    A, U, s, V = zo.make_matrix(50, "geometric", 1e7, 4)
    al, be = zo.bounds_estimate(A)
    assert al == pytest.approx(np.linalg.norm(A))
    assert be <= s.min() / al * (1 + 1e-12)
    assert be >= s.min() / al / (10 * math.sqrt(50))


def test_normest_vs_2norm():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((60, 40)) @ np.diag(np.geomspace(1, 1e-3, 40))
    e, cnt = zo.normest(A, 1e-10)
    assert e == pytest.approx(np.linalg.norm(A, 2), rel=1e-6)
    assert cnt >= 1


def test_metrics():
    n = 10
    Q = zo.haar(n, np.random.default_rng(1))
    assert zo.orthogonality(Q)[0] < 1e-15
    o2 = zo.orthogonality(2 * Q)
    assert o2[0] == pytest.approx(3.0 * math.sqrt(n) / math.sqrt(n))
    assert o2[1] == pytest.approx(3.0 * math.sqrt(n) / n)


# ---------------------------------------------------------------- mutation tests
def test_mutation_wrong_aj_sign_is_caught():
    """Flipping the sign convention of a_j breaks the partial-fraction identity."""
    con, rk = 1e6, 5
    coeff = zo.compute_coeff(con, EPS, rk)
    bad = coeff.copy()
    bad[2 * rk:] *= -1
    x = np.geomspace(1.0, con, 11)
    prod = _zolo_rational(x, coeff, rk) / x
    pf_bad = 1.0 - sum(bad[2 * rk + j] / (x * x + bad[2 * j]) for j in range(rk))
    assert not np.allclose(pf_bad, prod, rtol=1e-6)


def test_mutation_swapped_c_index_is_caught():
    """Using c_2j for c_2j-1 (an off-by-one in the coefficient index) breaks equioscillation."""
    con, rk = 1e4, 4
    coeff = zo.compute_coeff(con, EPS, rk)
    bad = coeff.copy()
    bad[0:2 * rk:2], bad[1:2 * rk:2] = coeff[1:2 * rk:2].copy(), coeff[0:2 * rk:2].copy()
    fone, con_next = zo.zolo_fone_con(con, coeff, rk)
    x = np.geomspace(1.0, con, 20001)
    g = _zolo_rational(x, bad, rk) / fone
    assert not (g.min() >= 1 - 1e-10 and g.max() <= con_next * (1 + 1e-10))


def test_mutation_unstructured_row_range_is_caught():
    """A structured QR that stops one row short (rows j..n+j-1) fails the LAPACK comparison."""
    n = 12
    rng = np.random.default_rng(2)
    S = np.zeros((2 * n, n))
    S[:n] = rng.standard_normal((n, n))
    S[n:] = 0.7 * np.eye(n)
    A = S.copy()
    for j in range(n):
        r1 = n + j  # one row short of the identity row n+j
        tau, beta, v = zo._larfg(A[j:r1, j].copy())
        A[j, j] = beta
        A[j + 1:r1, j] = v[1:]
        if j + 1 < n:
            w = v @ A[j:r1, j + 1:]
            A[j:r1, j + 1:] -= tau * np.outer(v, w)
    Rl = np.linalg.qr(S)[1]
    assert not np.allclose(np.abs(np.triu(A[:n])), np.abs(Rl), rtol=1e-6)


def test_mutation_qdwh_weight_is_caught():
    """A QDWH weight that drops the 1/3 power overshoots the [l, 1] bound."""
    ell = 1e-3
    x2 = ell * ell
    ww = 4.0 * (1.0 - x2) / (x2 * x2)  # missing cube root
    s1 = math.sqrt(1 + ww)
    a = s1 + math.sqrt(abs(8 - 4 * ww + 8 * (2 - x2) / (x2 * s1))) / 2
    b = (a - 1) ** 2 / 4
    c = a + b - 1
    x = np.geomspace(ell, 1.0, 10001)
    g = x * (a + b * x * x) / (1 + c * x * x)
    assert g.max() > 1 + 1e-6 or abs(a - zo.hfunc(ell)) > 1e-3


def _scalar_steps(beta, r, nsteps, npts=400):
    """Singular values of X0 = A / (alpha beta) filling [1, 1/beta], mapped by nsteps Zolotarev steps."""
    con = 1.0 / beta
    x = np.geomspace(1.0, con, npts)
    for _ in range(nsteps):
        coeff = zo.compute_coeff(con, zo.DLAMCH_P, r)
        a = coeff[2 * r:]
        c_odd = coeff[0:2 * r:2]
        fone, con = zo.zolo_fone_con(con, coeff, r)
        x = (x - sum(a[j] * x / (x * x + c_odd[j]) for j in range(r))) / fone
    return x


@pytest.mark.parametrize("beta", [0.3, 1e-2, 3e-4, 1e-6, 2e-9, 1e-11])
def test_batched_plan_converges_exactly_at_last_step(beta):
    r, sched = zo.batched_plan(beta)
    assert r == zo.choosem(1.0 / beta)
    assert 1 <= len(sched) <= 3 and set(sched) <= {"Q", "C"}
    assert sched[0] == ("Q" if beta < 1e-2 else sched[0])
    x = _scalar_steps(beta, r, len(sched))
    assert np.max(np.abs(x - 1.0)) <= 2 * zo.BATCH_CONVERGED
    x_short = _scalar_steps(beta, r, len(sched) - 1)
    assert np.max(np.abs(x_short - 1.0)) > zo.BATCH_CONVERGED


def test_zolo_svd_closed_form():
    n = 60
    rng = np.random.default_rng(17)
    U0 = zo.haar(n, rng)
    V0 = zo.haar(n, rng)
    s = 1.0 + 0.5 * np.sin(np.arange(1, n + 1)) + 1e-3 * rng.standard_normal(n)
    A = (U0 * s) @ V0.T
    S, U, V = zo.zolo_svd(A)
    assert np.allclose(S, np.sort(s), atol=1e-13)
    assert np.linalg.norm(A - (U * S) @ V.T) / np.linalg.norm(A) < 1e-14
    assert np.linalg.norm(U.T @ U - np.eye(n)) < 1e-13 and np.linalg.norm(V.T @ V - np.eye(n)) < 1e-13


def _near_singular(n=120, delta=2e-14, seed=23):
    rng = np.random.default_rng(seed)
    A = (zo.haar(n, rng) * np.geomspace(1.0, 1e-3, n)) @ zo.haar(n, rng).T
    A[:, 1] = A[:, 0]
    A[7, 0] = 0.0
    A[7, 1] = delta
    return A


def test_pzolopd2_restart_closed_form():
    """CON collapses to 1 after the first step (beta ~ 1e-16) while ERR is large:
    PZOLOPD2 restarts from the iterate and reaches the SVD polar factor."""
    A = _near_singular()
    al, be = zo.bounds_estimate(A)
    assert 1.0 / be > 5e15
    rk = zo.choosem(1.0 / be)
    U2, log2 = zo.pzolopd2(A, al, be, rk)
    assert log2.restarts >= 1
    W, s, Vt = np.linalg.svd(A)
    assert np.linalg.norm(U2 - W @ Vt) / np.sqrt(A.shape[0]) < 1e-12
    assert zo.orthogonality(U2)[0] < 1e-14


# ---------------------------------------------------------------- spectral division
def _sym_case(n, seed, sigma_frac=0.4):
    """Symmetric A = W diag(lam) W^T with a gap around sigma: lam in [-1, 1]
    with no eigenvalue within 0.05 of sigma."""
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(-1.0, 1.0, n))
    sigma = float(np.quantile(lam, sigma_frac))
    lam = np.where(np.abs(lam - sigma) < 0.05, lam + np.sign(lam - sigma + 1e-300) * 0.05, lam)
    A, W = zo.symmetric_matrix(n, lam, seed + 1)
    return A, np.sort(lam), sigma


@pytest.mark.parametrize("n,seed,frac,method", [(40, 1, 0.3, "zolo"), (90, 2, 0.6, "zolo"), (64, 3, 0.5, "qdwh")])
def test_eig_below_vs_eigvalsh(n, seed, frac, method):
    A, lam, sigma = _sym_case(n, seed, frac)
    w, V, log = zo.eig_below(A, sigma, method=method, seed=seed)
    exact = np.linalg.eigvalsh(A)
    below = exact[exact < sigma]
    assert log.rk == below.size and V.shape == (n, below.size)
    assert np.allclose(w, below, atol=1e-13)
    assert np.linalg.norm(A @ V - V * w) / np.linalg.norm(A) < 1e-13
    assert np.linalg.norm(V.T @ V - np.eye(below.size)) < 1e-13
    # span(V) = the exact invariant subspace: sin of the largest principal angle
    Wx = np.linalg.eigh(A)[1][:, exact < sigma]
    sines = np.linalg.svd(V - Wx @ (Wx.T @ V), compute_uv=False)
    assert sines.max() < 1e-12
    assert log.err[-1] <= 4.0 * 10.0 * EPS / 2.0 * 50


def test_pqdwhsubit_outputs():
    """On exit Q = V orthogonal and E = V^T A V has its off-diagonal block at
    the accepted level; Rk = ||C||_F^2 exactly for an exact sign function."""
    n = 50
    A, lam, sigma = _sym_case(n, 7, 0.5)
    W = np.linalg.eigh(A)[1]
    ev = np.linalg.eigvalsh(A)
    Q = (W * np.sign(ev - sigma)) @ W.T        # exact polar factor of A - sigma I
    V, E, log = zo.pqdwhsubit(A, float(np.linalg.norm(A)), Q, np.random.default_rng(0))
    assert log.rk == int((ev < sigma).sum())
    assert np.linalg.norm(V.T @ V - np.eye(n)) < 1e-13
    assert np.linalg.norm(E - V.T @ A @ V) < 1e-13
    # one iteration leaves err at a few eps (the reference's 4 TOL1 = 20 eps is
    # tight), the second always passes it
    assert log.iters <= 2 and log.err[0] < 1e-13 and log.err[-1] <= 4.0 * 10.0 * EPS / 2.0


def test_mutation_projector_sign_is_caught():
    """(Q + I)/2 projects onto the eigenvalues above sigma: the count pin fails."""
    n = 40
    A, lam, sigma = _sym_case(n, 11, 0.3)
    W = np.linalg.eigh(A)[1]
    ev = np.linalg.eigvalsh(A)
    Q = (W * np.sign(ev - sigma)) @ W.T
    _, _, log = zo.pqdwhsubit(A, float(np.linalg.norm(A)), Q, np.random.default_rng(0), projector_sign=+1.0)
    assert log.rk != int((ev < sigma).sum())


def test_eig_below_no_eigenvalue_below():
    n = 30
    A, lam, _ = _sym_case(n, 5, 0.5)
    w, V, log = zo.eig_below(A, float(lam[0]) - 0.5)
    assert log.rk == 0 and w.size == 0 and V.shape == (n, 0)


def test_svd_residual_metric():
    """check_routines.f90:70-86: zero (to rounding) for an exact SVD, and a
    change of delta in one singular value gives exactly delta / max(S)."""
    rng = np.random.default_rng(4)
    n = 30
    U, V = zo.haar(n, rng), zo.haar(n, rng)
    s = np.geomspace(3.0, 0.1, n)
    A = (U * s) @ V.T
    assert zo.svd_residual(A, U, s, V) < 1e-14
    s2 = s.copy()
    s2[7] += 1e-3
    assert math.isclose(zo.svd_residual(A, U, s2, V), 1e-3 / 3.0, rel_tol=1e-9)
    # the residual divides by the largest singular value, not the Frobenius norm
    assert math.isclose(zo.svd_residual(2 * A, U, 2 * s2, V), 1e-3 / 3.0, rel_tol=1e-9)
