"""CPU oracle: an FP64 numpy restatement of the reference Zolo-PD / QDWH-PD path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path (the HIP library and
its Fortran/Python front ends) never calls into it.

Every function cites the reference routine it restates.  The reference is
Fortran + ScaLAPACK (MKL); here a 1x1 process grid is assumed, so every
distributed PBLAS/ScaLAPACK call becomes the corresponding dense operation.

How each function is pinned (checked by ``tests/test_oracle.py``):

* ``mellipke`` / ``mellipj`` / ``compute_coeff`` -- against mpmath's ``ellipk``
  and ``ellipfun`` at 40 digits and against the Zolotarev equioscillation
  property of the resulting rational function.
* ``choosem`` -- the table itself (boundary values).
* ``hfunc`` -- the QDWH closed form (Nakatsukasa-Bai-Gygi 2010) and the
  fixed point a(1) = 3.
* ``structured_geqrf`` / ``structured_orgqr`` -- against LAPACK dgeqrf/dorgqr
  (numpy) on the full stacked matrix, and the invariants Q^T Q = I, QR = S.
* ``zoloqr_step`` / ``zolochol_step`` -- against X (X^T X + c I)^{-1} computed
  by a direct LAPACK solve, and against each other; ``zoloqr_step_lapack``
  (and ``pzolopd1(lapack_qr=True)``) against ``zoloqr_step`` (same schedule,
  same polar factor).
* ``pzolopd1`` / ``pqdwhfacs`` -- against the closed form U V^T of the
  generator A = U diag(s) V^T and against scipy.linalg.polar.
* ``pzolopd2`` -- the same closed form, on a numerically singular matrix
  where CON collapses to 1 after the first step and the restart is taken.
* ``zolo_svd`` -- A = U diag(S) V^T with S the generator's singular values
  (closed form), U and V orthogonal.
* ``batched_plan`` -- by applying the planned steps to diagonal matrices whose
  singular values fill [beta, 1]: the result is orthogonal to the FP32 unit
  roundoff after the last step and not before it.
* ``lacn2_norm1_inv`` -- against LAPACK dgecon (scipy) and the exact
  ||X^{-1}||_1 on small matrices; ``lacn2_norm1_inv_r`` / ``bounds_from_qr``
  against LAPACK dtrcon on R, the exact ||R^{-1}||_1, and beta <= sigma_min.
* ``normest`` -- against numpy's exact 2-norm.
* ``pqdwhsubit`` / ``zoloeig`` / ``eig_below`` -- against LAPACK's dsyevd
  (numpy.linalg.eigvalsh): the count of eigenvalues below sigma, those
  eigenvalues, the residual ||A V - V diag(w)||, V^T V = I, and the angle
  between span(V) and the exact invariant subspace; a plant of the wrong
  projector sign ((Q + I)/2) fails the count.
* ``mm_dense`` -- against scipy.io.mmread on files scipy.io.mmwrite wrote
  (symmetric, general, integer, pattern), where prepare_real_GEM's A + A^T and
  A(i,i)/2 - sigma must give scipy's full matrix with the diagonal shifted; a
  plant of the diagonal without the halving fails it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

# LAPACK DLAMCH('Precision') = eps*radix with eps = 2^-53, i.e. 2^-52.
DLAMCH_P = np.finfo(np.float64).eps


# ---------------------------------------------------------------------------
# Scalar math: Auxil_Zolo.f90, hfunc.f90
# ---------------------------------------------------------------------------

def mellipke(alpha: float, tol: float) -> tuple[float, float]:
    """Complete elliptic integrals K, E by the AGM (Auxil_Zolo.f90:36-71)."""
    m = math.sin(alpha) ** 2
    a0, b0, s0, i1, mm = 1.0, math.cos(alpha), m, 0, 1.0
    a1 = a0
    while mm > tol:
        a1 = (a0 + b0) / 2.0
        b1 = math.sqrt(a0 * b0)
        c1 = (a0 - b0) / 2.0
        i1 += 1
        w1 = (2.0 ** i1) * c1 * c1
        mm = w1
        s0 += w1
        a0, b0 = a1, b1
    k = math.pi / (2.0 * a1)
    e = k * (1.0 - s0 / 2.0)
    return k, e


def mellipj(u: float, alpha: float, tol: float) -> tuple[float, float, float]:
    """Jacobi sn, cn, dn by AGM + descending recurrence (Auxil_Zolo.f90:73-128)."""
    m = math.sin(alpha) ** 2
    mchunk = 1000
    a = [1.0]
    b = [math.cos(alpha)]
    c = [math.sin(alpha)]
    nin = 0  # 1-based index of the last c above tol, as in the Fortran
    i = 1
    while abs(c[i - 1]) > tol and i < mchunk:
        i += 1
        a.append((a[i - 2] + b[i - 2]) * 0.5)
        b.append(math.sqrt(a[i - 2] * b[i - 2]))
        c.append((a[i - 2] - b[i - 2]) * 0.5)
        if abs(c[i - 1]) <= tol and abs(c[i - 2]) > tol:
            nin = i - 1
    phin = [0.0] * (i + 1)
    phin[1] = u
    phin[i] = (2.0 ** nin) * a[i - 1] * u
    while i > 1:
        i -= 1
        phin[i] = phin[i + 1]
        if nin >= i:
            phin[i] = 0.5 * (math.asin(c[i] * math.sin(phin[i + 1]) / a[i]) + phin[i + 1])
    sn = math.sin(phin[1])
    cn = math.cos(phin[1])
    dn = math.sqrt(1.0 - m * sn * sn)
    return sn, cn, dn


def choosem(con: float) -> int:
    """Zolotarev degree rule r(kappa) (Auxil_Zolo.f90:130-163)."""
    if con < 1.001:
        return 2
    if con <= 1.01:
        return 3
    if con <= 1.1:
        return 4
    if con <= 1.2:
        return 5
    if con <= 1.5:
        return 6
    if con <= 2:
        return 8
    if con < 6.5:
        return 2
    if con < 180:
        return 3
    if con < 1.5e4:
        return 4
    if con < 2e6:
        return 5
    if con < 1e9:
        return 6
    if con < 3e12:
        return 7
    return 8


def compute_coeff(con: float, tol: float, rk: int) -> np.ndarray:
    """Zolotarev coefficients (Auxil_Zolo.f90:165-200).

    Returns coeff[0:3r]: coeff[0:2r] = c_1..c_2r, coeff[2r:3r] = a_1..a_r
    (0-based storage of the Fortran coeff(1:3r)).
    """
    kappa = 1.0 / con
    alpha = math.acos(kappa)
    kk, _ = mellipke(alpha, tol)
    coeff = np.zeros(3 * rk)
    rk2 = 2 * rk
    for i in range(1, rk2 + 1):
        t = i * kk / (rk2 + 1.0)
        sn, cn, _ = mellipj(t, alpha, tol)
        coeff[i - 1] = sn * sn / (cn * cn)
    for i in range(1, rk + 1):
        v = 1.0
        for j in range(1, rk + 1):
            v *= coeff[2 * i - 2] - coeff[2 * j - 1]
            if i != j:
                v /= coeff[2 * i - 2] - coeff[2 * j - 2]
        coeff[rk2 + i - 1] = v
    return coeff


def zolo_fone_con(con: float, coeff: np.ndarray, rk: int) -> tuple[float, float]:
    """FONE and the next CON (pzolopd1.f90:459-472)."""
    fone = 1.0
    for i in range(1, rk + 1):
        fone *= (1.0 + coeff[2 * i - 1]) / (1.0 + coeff[2 * i - 2])
    fcon = con
    for i in range(1, rk + 1):
        fcon *= (con * con + coeff[2 * i - 1]) / (con * con + coeff[2 * i - 2])
    return fone, max(fcon / fone, 1.0)


def hfunc(x: float) -> float:
    """QDWH weight a(l) (hfunc.f90:27-39)."""
    tol = 30 * DLAMCH_P
    x2 = x * x
    if abs(x - 1.0) <= tol:
        ww = 0.0
    else:
        tmp = 4.0 * (1.0 - x2) / (x2 * x2)
        ww = math.exp(math.log(tmp) / 3.0)
    sum1 = math.sqrt(1.0 + ww)
    h = math.sqrt(8.0 - 4.0 * ww + 8.0 * (2.0 - x2) / (x2 * sum1))
    return sum1 + h / 2.0


# ---------------------------------------------------------------------------
# Structured Householder QR of [X; s I]: mpdgeqrf.f / mpdorgqr.f
# ---------------------------------------------------------------------------

def _larfg(x: np.ndarray) -> tuple[float, float, np.ndarray]:
    """LAPACK dlarfg: H x = beta e1, H = I - tau v v^T, v[0] = 1."""
    alpha = x[0]
    xnorm = np.linalg.norm(x[1:])
    if xnorm == 0.0:
        v = np.zeros_like(x)
        v[0] = 1.0
        return 0.0, alpha, v
    beta = -math.copysign(math.hypot(alpha, xnorm), alpha)
    tau = (beta - alpha) / beta
    v = x / (alpha - beta)
    v[0] = 1.0
    return tau, beta, v


def structured_geqrf(S: np.ndarray, n: int) -> tuple[np.ndarray, np.ndarray]:
    """Householder QR of the 2n x n stack [X; s I] (mpdgeqrf.f:262-313).

    Column j (0-based) only touches rows j .. n+j: rows below n+j are still
    zero in columns >= j (the identity block fills in one row per column).
    Returns the compact form (R in the upper triangle, reflectors below) and tau.
    """
    A = np.array(S, dtype=np.float64, copy=True)
    tau = np.zeros(n)
    for j in range(n):
        r1 = n + j + 1  # exclusive end of the active rows
        tau[j], beta, v = _larfg(A[j:r1, j].copy())
        A[j, j] = beta
        A[j + 1:r1, j] = v[1:]
        if j + 1 < n and tau[j] != 0.0:
            w = v @ A[j:r1, j + 1:]
            A[j:r1, j + 1:] -= tau[j] * np.outer(v, w)
    return A, tau


def structured_orgqr(F: np.ndarray, tau: np.ndarray, n: int) -> np.ndarray:
    """Explicit Q (2n x n) from the structured reflectors (mpdorgqr.f:253-337)."""
    Q = np.zeros((2 * n, n))
    Q[:n, :n] = np.eye(n)
    for j in range(n - 1, -1, -1):
        r1 = n + j + 1
        v = np.zeros(r1 - j)
        v[0] = 1.0
        v[1:] = F[j + 1:r1, j]
        w = v @ Q[j:r1, j:]
        Q[j:r1, j:] -= tau[j] * np.outer(v, w)
    return Q


# ---------------------------------------------------------------------------
# Per-shift kernels: zoloqr.f90, zolochol.f90, FormX2.f90
# ---------------------------------------------------------------------------

def zoloqr_step(X: np.ndarray, coeff: np.ndarray, rk: int, ip: int) -> np.ndarray:
    """Z_ip = -a_ip / sqrt(c_{2ip-1}) * Q1 Q2^T, [X; sqrt(c) I] = QR (zoloqr.f90:82-112).

    ip is 1-based as in the reference.
    """
    n = X.shape[0]
    c = coeff[2 * ip - 2]
    S = np.zeros((2 * n, n))
    S[:n] = X
    S[n:] = math.sqrt(c) * np.eye(n)
    F, tau = structured_geqrf(S, n)
    Q = structured_orgqr(F, tau, n)
    akk = -coeff[2 * rk + ip - 1] / math.sqrt(c)
    return akk * (Q[:n] @ Q[n:].T)


def zoloqr_step_lapack(X: np.ndarray, coeff: np.ndarray, rk: int, ip: int) -> np.ndarray:
    """zoloqr_step with LAPACK's blocked dgeqrf + dorgqr on the whole stack.

    The structure of [X; sqrt(c) I] is only exploited by the column loop of
    structured_geqrf; LAPACK's QR of the same stack gives the same Q up to
    column signs, which cancel in Q1 Q2^T.  This is the form the CPU baseline
    times (level-3 BLAS, all host cores).
    """
    n = X.shape[0]
    c = coeff[2 * ip - 2]
    S = np.empty((2 * n, n))
    S[:n] = X
    S[n:] = 0.0
    S[n:][np.diag_indices(n)] = math.sqrt(c)
    Q = sla.qr(S, mode="economic", overwrite_a=True, check_finite=False)[0]
    akk = -coeff[2 * rk + ip - 1] / math.sqrt(c)
    return akk * (Q[:n] @ Q[n:].T)


def zolochol_step(X: np.ndarray, coeff: np.ndarray, rk: int, ip: int, sym: str = "N") -> np.ndarray:
    """Z_ip = -a_ip X (X^T X + c_{2ip-1} I)^{-1} (zolochol.f90:87-113).

    SYM='S' solves with X itself as the right-hand side (zolochol.f90:99-100),
    which equals the 'N' branch only when X is symmetric.
    """
    n = X.shape[0]
    W = X.T @ X
    W[np.diag_indices(n)] += coeff[2 * ip - 2]
    cf = sla.cho_factor(W, lower=False)
    if sym.upper().startswith("S"):
        Z = sla.cho_solve(cf, X)
    else:
        Z = sla.cho_solve(cf, X.T).T
    return -coeff[2 * rk + ip - 1] * Z


def formx2(X: np.ndarray, Z: np.ndarray, fone: float) -> tuple[np.ndarray, float]:
    """X+ = (X + Z)/FONE, ERR = ||X+ - X||_F (FormX2.f90:73-89).

    PDLASCL('General', FONE, ONE) at FormX2.f90:74 scales by MUL = CTO/CFROM
    = 1/FONE (xLASCL's single-step case), so X+ is (X + Z) * (1/FONE)."""
    Xp = (X + Z) * (1.0 / fone)
    return Xp, float(np.linalg.norm(Xp - X))


# ---------------------------------------------------------------------------
# Bounds: test_zolodrv1.f90:179-189 (alpha, beta), LAPACK dlacn2 / dgecon
# ---------------------------------------------------------------------------

def lacn2_norm1_inv(lu_piv) -> tuple[float, int]:
    """Estimate ||X^{-1}||_1 from an LU factorisation (Higham 1988; LAPACK dlacn2).

    This is the estimator behind PDGECON('1') used at test_zolodrv1.f90:186 and
    pqdwhfacs.f90:340.  As in xGECON, the solves use inv(U) inv(L) without the
    row permutation (a column permutation of X^{-1} leaves its 1-norm
    unchanged).  The reverse-communication loop of dlacn2 is unrolled here.
    Returns (estimate, number of solves).
    """
    lu, _ = lu_piv
    n = lu.shape[0]

    def solve(b, trans):
        if trans == 0:   # inv(U) inv(L) b
            y = sla.solve_triangular(lu, b, lower=True, unit_diagonal=True)
            return sla.solve_triangular(lu, y, lower=False)
        y = sla.solve_triangular(lu, b, lower=False, trans=1)   # inv(U^T) b
        return sla.solve_triangular(lu, y, lower=True, unit_diagonal=True, trans=1)

    return _lacn2(solve, n)


def lacn2_norm1_inv_r(R: np.ndarray) -> tuple[float, int]:
    """The same estimator on the upper triangle of R (DTRCON('1', 'U', 'N')'s use
    of dlacn2): an estimate of ||R^{-1}||_1.  Returns (estimate, solves)."""
    Ru = np.triu(R)

    def solve(b, trans):
        return sla.solve_triangular(Ru, b, lower=False, trans=trans)

    return _lacn2(solve, Ru.shape[0])


def _lacn2(solve, n: int) -> tuple[float, int]:
    """dlacn2's reverse-communication loop, unrolled; solve(b, trans) applies
    M^{-1} (trans 0) or M^{-T} (trans 1)."""
    nsolve = 0
    x = solve(np.full(n, 1.0 / n), 0); nsolve += 1
    if n == 1:
        return abs(x[0]), nsolve
    est = float(np.abs(x).sum())
    sgn = np.where(x >= 0, 1.0, -1.0)
    x = solve(sgn, 1); nsolve += 1
    j = int(np.argmax(np.abs(x)))
    it = 2
    while True:
        e = np.zeros(n); e[j] = 1.0
        x = solve(e, 0); nsolve += 1
        estold = est
        est = float(np.abs(x).sum())
        newsgn = np.where(x >= 0, 1.0, -1.0)
        if np.array_equal(newsgn, sgn) or est <= estold:
            break
        sgn = newsgn
        x = solve(sgn, 1); nsolve += 1
        jlast = j
        j = int(np.argmax(np.abs(x)))
        if x[jlast] != abs(x[j]) and it < 5:
            it += 1
            continue
        break
    altsgn = np.array([(1.0 if i % 2 == 0 else -1.0) * (1.0 + i / (n - 1)) for i in range(n)])
    x = solve(altsgn, 0); nsolve += 1
    temp = 2.0 * float(np.abs(x).sum()) / (3.0 * n)
    if temp > est:
        est = temp
    return est, nsolve


def bounds_estimate(A: np.ndarray, convention: str = "zolo") -> tuple[float, float]:
    """alpha = ||A||_F and the lower bound beta (test_zolodrv1.f90:179-189).

    convention='zolo':  beta = rcond / sqrt(n)                 (test_zolodrv1.f90:189)
    convention='qdwh':  beta = ||A/alpha||_1 * rcond / sqrt(n)  (pqdwhfacs.f90:344)
    rcond = 1 / (1 * ||X^{-1}||_1 estimate), X = A/alpha (PDGECON with ANORM=1).
    """
    n = A.shape[0]
    alpha = float(np.linalg.norm(A))
    X = A / alpha
    lu_piv = sla.lu_factor(X)
    ainvnm, _ = lacn2_norm1_inv(lu_piv)
    rcond = 1.0 / ainvnm
    beta = rcond / math.sqrt(n)
    if convention == "qdwh":
        beta *= float(np.abs(X).sum(axis=0).max())
    return alpha, beta


def bounds_from_qr(A: np.ndarray) -> tuple[float, float]:
    """alpha = ||A||_F and beta = 1 / (||R^{-1}||_1 estimate * sqrt(n)) for
    A/alpha = Q R: test_zolodrv1.f90:179-189's bound with the estimate taken on
    R (||R^{-1}||_2 = ||(A/alpha)^{-1}||_2 <= sqrt(n) ||R^{-1}||_1, so beta is a
    lower bound on sigma_min(A/alpha) exactly as the LU form)."""
    n = A.shape[0]
    alpha = float(np.linalg.norm(A))
    R = np.linalg.qr(A / alpha, mode="r")
    est, _ = lacn2_norm1_inv_r(R)
    return alpha, min(1.0, 1.0 / (est * math.sqrt(n)))


def normest(A: np.ndarray, tol: float = 1e-6, maxiter: int = 100) -> tuple[float, int]:
    """2-norm estimate by power iteration, MATLAB normest (pdnormest.f90:43-80).

    The reference body multiplies by A twice (pdnormest.f90:68); the second
    product of normest is A^T, which is what is restated here.
    """
    x = np.abs(A).sum(axis=0)
    e = float(np.linalg.norm(x))
    if e == 0:
        return 0.0, 0
    x = x / e
    e0 = 0.0
    cnt = 0
    while abs(e - e0) > tol * e and cnt < maxiter:
        e0 = e
        sx = A @ x
        nsx = float(np.linalg.norm(sx))
        x = A.T @ sx
        nx = float(np.linalg.norm(x))
        e = nx / nsx
        x = x / nx
        cnt += 1
    return e, cnt


# ---------------------------------------------------------------------------
# Drivers: pzolopd1.f90, pqdwhfacs.f90
# ---------------------------------------------------------------------------

@dataclass
class ZoloLog:
    schedule: list = field(default_factory=list)   # 'QR' / 'CHOL' per iteration
    err: list = field(default_factory=list)        # ERR after each FormX2
    con: list = field(default_factory=list)        # CON at the start of each iteration
    maxc: list = field(default_factory=list)
    tol2: float = 0.0
    iters: int = 0
    restarts: int = 0


def pzolopd1(A: np.ndarray, alpha: float, beta: float, rk: int, sym: str = "N",
             max_it: int = 6, lapack_qr: bool = False) -> tuple[np.ndarray, ZoloLog]:
    """Zolo-PD (pzolopd1.f90:280-487) on one process, r shifts done in sequence.

    Returns the polar factor U_p and the iteration log.  lapack_qr = True
    takes the QR steps with zoloqr_step_lapack (the timed CPU baseline).
    """
    qr_step = zoloqr_step_lapack if lapack_qr else zoloqr_step
    n = A.shape[0]
    X = A / (alpha * beta)
    con = 1.0 / beta
    tol = DLAMCH_P
    tol1 = 10.0 * DLAMCH_P / 2.0
    tol2 = math.exp(math.log(tol1) / (2.0 * rk + 1.0))
    log = ZoloLog(tol2=tol2)
    it, err = 0, 1.0
    while err > tol2 and it < max_it:
        coeff = compute_coeff(con, tol, rk)
        maxc = float(coeff[:2 * rk].max())
        log.con.append(con)
        log.maxc.append(maxc)
        Z = np.zeros_like(X)
        if (it == 0 or con > 1e3) and maxc > 3e2:
            log.schedule.append("QR")
            for ip in range(1, rk + 1):
                Z += qr_step(X, coeff, rk, ip)
        else:
            log.schedule.append("CHOL")
            for ip in range(1, rk + 1):
                Z += zolochol_step(X, coeff, rk, ip, sym)
        fone, con = zolo_fone_con(con, coeff, rk)
        X, err = formx2(X, Z, fone)
        log.err.append(err)
        it += 1
    log.iters = it
    return X, log


BATCH_CONVERGED = 6e-8   # CON - 1 at which an FP32 iterate has converged
BATCH_F64_QR_CON = 1e8   # a first QR step above this CON runs in FP64 on the device


def batched_plan(beta: float, rk: int = 0, max_it: int = 6) -> tuple[int, str]:
    """Schedule of one matrix of the batched FP32 path (pzolopd1.f90:286-470).

    CON follows a scalar recurrence that only depends on beta, so the whole
    schedule is known in advance: r = choosem(1/beta), step k is a QR step when
    (k == 0 or CON > 1e3) and max c > 300, else a Cholesky step, and the matrix
    stops after the step whose CON - 1 is at most BATCH_CONVERGED.  Returns
    (r, schedule string of 'Q' / 'C').  (On the device a first QR step at
    CON > BATCH_F64_QR_CON runs in FP64; the schedule is the same.)
    """
    con = 1.0 / min(beta, 1.0)
    r = rk if rk > 0 else choosem(con)
    sched = ""
    if con - 1.0 <= BATCH_CONVERGED:
        return r, sched
    for it in range(max_it):
        coeff = compute_coeff(con, DLAMCH_P, r)
        maxc = float(coeff[:2 * r].max())
        sched += "Q" if (it == 0 or con > 1e3) and maxc > 3e2 else "C"
        _, con = zolo_fone_con(con, coeff, r)
        if con - 1.0 <= BATCH_CONVERGED:
            break
    return r, sched


def pzolopd2(A: np.ndarray, alpha: float, beta: float, rk: int, sym: str = "N",
             max_it: int = 12) -> tuple[np.ndarray, ZoloLog]:
    """PZOLOPD2 (pzolopd2.f90:386-487): PZOLOPD1 with IT < 12 and a restart.

    When the CON recurrence reaches exactly 1 while ERR is still above TOL2 the
    reference calls FormXTwo (pzolopd2.f90:480-486), which it never defines;
    restated here as: re-estimate alpha, beta of the current iterate
    (test_zolodrv1.f90:179-189), X <- X/(alpha beta), CON <- 1/beta, continue.
    log.restarts counts them.
    """
    n = A.shape[0]
    X = A / (alpha * beta)
    con = 1.0 / beta
    tol1 = 10.0 * DLAMCH_P / 2.0
    tol2 = math.exp(math.log(tol1) / (2.0 * rk + 1.0))
    log = ZoloLog(tol2=tol2)
    it, err = 0, 1.0
    while err > tol2 and it < max_it:
        coeff = compute_coeff(con, DLAMCH_P, rk)
        maxc = float(coeff[:2 * rk].max())
        log.con.append(con)
        log.maxc.append(maxc)
        Z = np.zeros_like(X)
        if (it == 0 or con > 1e3) and maxc > 3e2:
            log.schedule.append("QR")
            for ip in range(1, rk + 1):
                Z += zoloqr_step(X, coeff, rk, ip)
        else:
            log.schedule.append("CHOL")
            for ip in range(1, rk + 1):
                Z += zolochol_step(X, coeff, rk, ip, sym)
        fone, con = zolo_fone_con(con, coeff, rk)
        X, err = formx2(X, Z, fone)
        log.err.append(err)
        it += 1
        if con == 1.0 and err > tol2 and it < max_it:
            ax, bx = bounds_estimate(X)
            if 0.0 < bx < 1.0:
                X = X / (ax * bx)
                con = 1.0 / bx
                log.restarts += 1
    log.iters = it
    return X, log


@dataclass
class QdwhLog:
    schedule: list = field(default_factory=list)
    err: list = field(default_factory=list)
    lowk: list = field(default_factory=list)
    alpha: float = 0.0
    beta: float = 0.0
    iters: int = 0


def pqdwhfacs(A: np.ndarray, sigma: float = 0.0, polar: bool = True, sym: str = "N"):
    """QDWH-PD (pqdwhfacs.f90:296-466). Returns (U_p, H or None, log)."""
    n = A.shape[0]
    tol1 = 10.0 * DLAMCH_P / 2.0
    tol2 = math.exp(math.log(tol1) / 3.0)
    Q = np.array(A, dtype=np.float64, copy=True)
    if sigma != 0.0:
        Q[np.diag_indices(n)] -= sigma
    alpha = float(np.linalg.norm(Q))
    Q = Q / alpha
    temp_beta = float(np.abs(Q).sum(axis=0).max())
    lu_piv = sla.lu_factor(Q)
    ainvnm, _ = lacn2_norm1_inv(lu_piv)
    beta = temp_beta * (1.0 / ainvnm) / math.sqrt(n)
    log = QdwhLog(alpha=alpha, beta=beta)
    it, lowk, err = 0, beta, 1.0
    err2 = abs(1.0 - lowk)
    while (it == 0 or err > tol2 or err2 > tol1) and it <= 12:
        ak = hfunc(lowk)
        bk = (ak - 1.0) ** 2 / 4.0
        ck = ak + bk - 1.0
        sck = math.sqrt(ck)
        lowk2 = lowk * lowk
        lowk = lowk * (ak + bk * lowk2) / (1.0 + ck * lowk2)
        if ck > 100:
            log.schedule.append("QR")
            S = np.zeros((2 * n, n))
            S[:n] = Q
            S[n:] = np.eye(n) / sck
            F, tau = structured_geqrf(S, n)
            QQ = structured_orgqr(F, tau, n)
            W2 = QQ[:n] @ QQ[n:].T
            X1 = (bk / ck) * Q + ((ak - bk / ck) / sck) * W2
        else:
            log.schedule.append("CHOL")
            W = ck * (Q.T @ Q)
            W[np.diag_indices(n)] += 1.0
            cf = sla.cho_factor(W, lower=False)
            if sym.upper().startswith("S"):
                W2 = sla.cho_solve(cf, Q)
            else:
                W2 = sla.cho_solve(cf, Q.T).T
            X1 = (bk / ck) * Q + (ak - bk / ck) * W2
        err = float(np.linalg.norm(Q - X1))
        it += 1
        err2 = abs(1.0 - lowk)
        log.err.append(err)
        log.lowk.append(lowk)
        Q = X1
    log.iters = it
    H = None
    if polar:
        H = polar_h(Q, A)
    return Q, H, log


def polar_h(U: np.ndarray, A: np.ndarray) -> np.ndarray:
    """H = (U^T A + A^T U)/2 (pqdwhfacs.f90:460-466, ZoloCHK1.f90:80-83)."""
    T = U.T @ A
    return 0.5 * (T + T.T)


def zolo_svd(A: np.ndarray, rk: int = 0):
    """Zolo-SVD (ZoloCHK1.f90:80-91): polar factor, H, eigendecomposition, U = U_p V.

    The eigensolver step (ELPA_PDSYEVD2, ZoloCHK1.f90:86) is LAPACK's dsyevd
    here (numpy.linalg.eigh); eigenvalues ascending as ELPA returns them.
    Returns (S ascending, U, V) with A = U diag(S) V^T.
    """
    alpha, beta = bounds_estimate(A)
    r = rk if rk > 0 else choosem(1.0 / beta)
    Up, _ = pzolopd1(A, alpha, beta, r)
    H = polar_h(Up, A)
    w, V = np.linalg.eigh(H)
    return w, Up @ V, V


# ---------------------------------------------------------------------------
# Spectral division: pqdwhsubit.f90, ZoloEig.f90, pqdwheig.f90
# ---------------------------------------------------------------------------

@dataclass
class SubitLog:
    rk: int = 0
    iters: int = 0
    err: list = field(default_factory=list)   # ||E(1:Rk, Rk+1:N)||_F / normA per iteration


def pqdwhsubit(A: np.ndarray, normA: float, Q: np.ndarray, rng: np.random.Generator,
               projector_sign: float = -1.0):
    """PQDWHSUBIT (pqdwhsubit.f90:223-316) on one process.

    Q is the polar factor of A - sigma I.  C = (Q - I)/2 (pqdwhsubit.f90:228-239),
    Rk = NINT(||C||_F^2) (241-243), oversampling pp = 3 (225, 244); random start
    block (259), QR + ORGQR(N, kk, kk) (266-267), Y = C X1 (269), QR +
    ORGQR(N, N, kk) (271-272), E = V^T A V (276-279), err on E(1:Rk, Rk+1:N)
    (282-285), accept at err <= 4 TOL1 (292) or after the second iteration
    (297), else restart from C V1 (305-311).  The start block is uniform
    [0, 1) as Fortran's random_number, from ``rng``.  ``projector_sign`` = +1
    plants the wrong sign ((Q + I)/2) for the tests.
    Returns (V, E, log).
    """
    n = A.shape[0]
    tol1 = 10.0 * DLAMCH_P / 2.0
    C = 0.5 * (Q + projector_sign * np.eye(n))
    rk = int(round(float(np.linalg.norm(C)) ** 2))
    rk = min(max(rk, 0), n)
    kk = min(rk + 3, n)
    log = SubitLog(rk=rk)
    X = rng.uniform(0.0, 1.0, (n, kk))
    it = 1
    while True:
        X1, _ = np.linalg.qr(X, mode="reduced")
        Y = C @ X1
        V, _ = np.linalg.qr(Y, mode="complete")
        E = V.T @ (A @ V)
        err = float(np.linalg.norm(E[:rk, rk:])) / normA
        log.err.append(err)
        if err <= 4.0 * tol1 or it == 2:
            break
        X = C @ V[:, :rk]
        kk = rk
        it += 1
    log.iters = it
    return V, E, log


def zoloeig(Q: np.ndarray, As: np.ndarray, rng: np.random.Generator):
    """ZoloEig (ZoloEig.f90:80-103): PQDWHSUBIT, then dsyevd on the leading
    Rk x Rk block of E (upper triangle, as PDSYEVD('V', 'U')), V1 VV1.
    Returns (w ascending, V (n x Rk), log)."""
    normA = float(np.linalg.norm(As))
    V, E, log = pqdwhsubit(As, normA, Q, rng)
    rk = log.rk
    if rk == 0:
        return np.zeros(0), np.zeros((As.shape[0], 0)), log
    w, VV = np.linalg.eigh(E[:rk, :rk], UPLO="U")
    return w, V[:, :rk] @ VV, log


def eig_below(A: np.ndarray, sigma: float, rk: int = 0, method: str = "zolo", seed: int = 0):
    """pqdwheig.f90:113-144: polar factor of A - sigma I, then ZoloEig.

    method 'zolo': Zolo-PD with the bounds of test_zolodrv1.f90 (sym 'S');
    'qdwh': pqdwhfacs with SIGMA as the reference driver calls it.
    Returns (w, V, log) -- the eigenpairs of A below sigma."""
    n = A.shape[0]
    if method == "qdwh":
        Q, _, _ = pqdwhfacs(A, sigma, polar=False, sym="S")
    else:
        B = np.array(A, dtype=np.float64, copy=True)
        B[np.diag_indices(n)] -= sigma
        alpha, beta = bounds_estimate(B)
        r = rk if rk > 0 else choosem(1.0 / beta)
        Q, _ = pzolopd1(B, alpha, beta, r, sym="S")
    return zoloeig(Q, A, np.random.default_rng(seed))


def mm_dense(path: str, sigma: float = 0.0, halve_diagonal: bool = True):
    """Dense matrix of a MatrixMarket coordinate file as prepare_real_GEM
    (prepare_matrix.f90:214-321) and prepare_real_symm_MM (:638-759) assemble
    it: entries stored in file order (A(IPQ) = rval(I), :276-283), then for
    'symmetric' files A <- A + A^T (PDTRAN, :292) and A(i,i) <- A(i,i)/2 -
    sigma (:293-302); other types as stored.  mmread (:1076-1399) supplies the
    entries; 'pattern' entries are 1 here.  ``halve_diagonal`` = False plants
    a mistake for the tests.  Returns (A, mtype)."""
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[0].split()
    symm = head[4].lower()
    field = head[3].lower()
    mtype = 0 if symm == "symmetric" else 1 if symm == "general" else 2
    k = 1
    while lines[k].startswith("%") or not lines[k].strip():
        k += 1
    rows, cols, nnz = (int(t) for t in lines[k].split()[:3])
    A = np.zeros((rows, cols))
    got = 0
    for ln in lines[k + 1:]:
        if got == nnz:
            break
        if not ln.strip() or ln.startswith("%"):
            continue
        t = ln.split()
        v = 1.0 if field == "pattern" else float(t[2].replace("D", "e").replace("d", "e"))
        A[int(t[0]) - 1, int(t[1]) - 1] = v
        got += 1
    if mtype == 0:
        A = A + A.T
        d = np.diag_indices(rows)
        A[d] = (A[d] / 2.0 if halve_diagonal else A[d]) - sigma
    return A, mtype


def symmetric_matrix(n: int, lam: np.ndarray, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """A = W diag(lam) W^T with Haar W (symmetrised); returns (A, W)."""
    W = haar(n, np.random.default_rng(seed))
    A = (W * lam) @ W.T
    return 0.5 * (A + A.T), W


# ---------------------------------------------------------------------------
# Metrics (check_routines.f90:142-153, ZoloCHK.f90:196-200) and generators
# ---------------------------------------------------------------------------

def backward_error(A, U, H) -> float:
    """||A - U H||_F / ||A||_F (relative form of ZoloCHK.f90:196-200)."""
    return float(np.linalg.norm(A - U @ H) / np.linalg.norm(A))


def orthogonality(U) -> tuple[float, float]:
    """(||U^T U - I||_F / sqrt(n), ||U^T U - I||_F / n) (check_routines.f90:142-153)."""
    n = U.shape[1]
    e = float(np.linalg.norm(U.T @ U - np.eye(n)))
    return e / math.sqrt(n), e / n


def svd_residual(A, U, S, V) -> float:
    """||A V - U diag(S)||_F / max(S) (CheckResidual_SVD, check_routines.f90:70-86;
    ZoloCHK1.f90:103-134)."""
    S = np.asarray(S)
    return float(np.linalg.norm(A @ V - U * S) / S.max())


def haar(n: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random orthogonal matrix: QR of an i.i.d. N(0,1) matrix, signs fixed."""
    G = rng.standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    return Q * np.sign(np.diag(R))


def spectrum(kind: str, n: int, rng: np.random.Generator, kappa: float = 1e6) -> np.ndarray:
    """Singular value profiles of the workloads (descending)."""
    if kind == "geometric":
        return kappa ** (-np.arange(n) / max(n - 1, 1))
    if kind == "loguniform":          # s_1 = 1, s_n = 1/kappa pinned, others log-uniform
        s = kappa ** (-rng.uniform(0.0, 1.0, n))
        s[0], s[-1] = 1.0, 1.0 / kappa
        return np.sort(s)[::-1]
    if kind == "cluster":             # geometric 1 -> 1/kappa, last 5% exactly 1/kappa
        m = n - n // 20
        s = np.full(n, 1.0 / kappa)
        s[:m] = kappa ** (-np.arange(m) / max(m - 1, 1))
        return s
    if kind == "sine":                # 1 + 0.5 sin(k) + 1e-3 N(0,1), k = 1..n
        k = np.arange(1, n + 1)
        return 1.0 + 0.5 * np.sin(k) + 1e-3 * rng.standard_normal(n)
    raise ValueError(kind)


def make_matrix(n: int, kind: str, kappa: float = 1e6, seed: int = 0):
    """A = U diag(s) V^T with Haar U, V; returns (A, U, s, V)."""
    rng = np.random.default_rng(seed)
    U = haar(n, rng)
    V = haar(n, rng)
    s = spectrum(kind, n, rng, kappa)
    return (U * s) @ V.T, U, s, V
