"""The five workloads at their full sizes on the GPU.

The CPU oracle cannot run n = 8192 or 16384 in seconds, so these sizes are
checked through properties that do not depend on n:
  * orthogonality ||U^T U - I||_F / sqrt(n) and backward error
    ||A - U H||_F / ||A||_F at the FP64 level (1e-13);
  * the eigenvalues of H (torch's eigensolver) against the generated singular
    values: H = V diag(s) V^T exactly for A = U diag(s) V^T, and Weyl's
    inequality bounds the difference by the backward error times ||A||;
  * the bounds alpha (Frobenius norm) and beta (the oracle's LU + 1-norm
    estimate on the same data, on the host), r = choosem(1/beta);
  * the CON recurrence and the QR/Cholesky rule of the oracle
    (pzolopd1.f90:405,459-472), which depend only on beta and r;
  * mixed against FP64 polar factors, within ten times the FP64 result's own
    distance to the generator's U V^T (the conditioning of the polar factor
    of the rounded A sets both);
  * for the batch of 4096 matrices, the oracle's scalar plan of every matrix.
QDWH at n = 1024 runs against the oracle itself.
"""
import math

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _errors(torch, A, dU, dH):
    n = A.shape[0]
    U, H = dU.T, dH.T
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    bwd = (torch.linalg.norm(A - U @ H) / torch.linalg.norm(A)).item()
    orth = (torch.linalg.norm(U.T @ U - eye) / math.sqrt(n)).item()
    return bwd, orth


def _scalar_steps(beta, rk, iters):
    """CON at the start of each step and the step kind, from beta alone (pzolopd1.f90:286,397-472)."""
    con = 1.0 / beta
    cons, kinds = [], ""
    for it in range(iters):
        coeff = zo.compute_coeff(con, zo.DLAMCH_P, rk)
        maxc = float(coeff[:2 * rk].max())
        cons.append(con)
        kinds += "Q" if (it == 0 or con > 1e3) and maxc > 3e2 else "C"
        _, con = zo.zolo_fone_con(con, coeff, rk)
    return cons, kinds


def _spectrum_err(torch, dH, s):
    lam = torch.linalg.eigvalsh(dH.T.contiguous())
    return torch.max(torch.abs(torch.sort(lam).values - torch.sort(s).values)).item()


def test_qdwh_n1024_kappa1e6_matches_oracle(zl, torch_gpu):
    n = 1024
    A, Ug, s, Vg = zo.make_matrix(n, "geometric", 1e6, 20260101)
    dU, dH, info = zl.pqdwhfacs_device(zl.to_device(A))
    torch_gpu.cuda.synchronize()
    U, H = zl.from_device(dU), zl.from_device(dH)
    Uo, Ho, lo = zo.pqdwhfacs(A)
    sched = "".join("Q" if x == "QR" else "C" for x in lo.schedule)
    assert info.schedule == sched and info.iters == lo.iters, (info.schedule, sched)
    assert info.beta == pytest.approx(lo.beta, rel=1e-8)
    be, orth = zo.backward_error(A, U, H), zo.orthogonality(U)[0]
    beo, ortho = zo.backward_error(A, Uo, Ho), zo.orthogonality(Uo)[0]
    assert be <= 10 * max(beo, 1e-15) and orth <= 10 * max(ortho, 1e-15), (be, beo, orth, ortho)
    d_o = np.linalg.norm(U - Uo) / math.sqrt(n)
    d_cf_o = np.linalg.norm(Uo - Ug @ Vg.T) / math.sqrt(n)
    assert d_o <= 10 * max(d_cf_o, 1e-14), (d_o, d_cf_o)


def test_zolo_f64_n8192_kappa1e12(zl, torch_gpu):
    from zolo import workloads
    torch = torch_gpu
    n = 8192
    A, dA, U0, s, V0 = workloads.square(n, "loguniform", 1e12, 20260102)
    alpha, beta = zl.bounds_device(dA)
    assert alpha == pytest.approx(torch.linalg.norm(A).item(), rel=1e-13)
    _, beta_o = zo.bounds_estimate(A.cpu().numpy())
    assert beta == pytest.approx(beta_o, rel=1e-6), (beta, beta_o)
    rk = zl.choosem(1.0 / beta)
    assert rk == zo.choosem(1.0 / beta) and 6 <= rk <= 8
    dU, dH, info = zl.pzolopd1_device(dA, alpha, beta, rk, precision="f64")
    torch.cuda.synchronize()
    cons, kinds = _scalar_steps(beta, rk, info.iters)
    assert info.schedule == kinds and info.schedule.startswith("QC"), (info.schedule, kinds)
    assert np.allclose(info.con, cons, rtol=1e-9)
    bwd, orth = _errors(torch, A, dU, dH)
    assert bwd <= TOL and orth <= TOL, (bwd, orth)
    assert _spectrum_err(torch, dH, s) <= TOL


@pytest.fixture(scope="module")
def c3(zl, torch_gpu):
    """n = 16384, s geometric 1 -> 1e-8 with the last 5 % at 1e-8, and its FP64 polar decomposition."""
    from zolo import workloads
    torch = torch_gpu
    A, dA, U0, s, V0 = workloads.square(16384, "cluster", 1e8, 20260103)
    alpha, beta = zl.bounds_device(dA)
    rk = zl.choosem(1.0 / beta)
    dU, dH, info = zl.pzolopd1_device(dA, alpha, beta, rk, precision="f64")
    torch.cuda.synchronize()
    d_cf = (torch.linalg.norm(dU.T - U0 @ V0.T) / math.sqrt(A.shape[0])).item()
    del U0, V0
    yield {"A": A, "dA": dA, "s": s, "alpha": alpha, "beta": beta, "rk": rk, "U": dU, "H": dH, "info": info,
           "d_cf": d_cf}
    del A, dA, dU, dH
    torch.cuda.empty_cache()


def test_zolo_f64_n16384_cluster(zl, torch_gpu, c3):
    torch = torch_gpu
    info = c3["info"]
    cons, kinds = _scalar_steps(c3["beta"], c3["rk"], info.iters)
    assert info.schedule == kinds and info.schedule[0] == "Q", (info.schedule, kinds)
    assert np.allclose(info.con, cons, rtol=1e-9)
    bwd, orth = _errors(torch, c3["A"], c3["U"], c3["H"])
    assert bwd <= TOL and orth <= TOL, (bwd, orth)


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_zolo_mixed_n16384_cluster(zl, torch_gpu, c3, prec):
    torch = torch_gpu
    dU, dH, info = zl.pzolopd1_device(c3["dA"], c3["alpha"], c3["beta"], c3["rk"], precision=prec)
    torch.cuda.synchronize()
    # FP64 first (QR) step, low-precision Cholesky steps, FP64 polish
    assert info.schedule[0] == "Q" and set(info.schedule[1:]) <= {"c"} and info.polish_passes >= 1, info.schedule
    bwd, orth = _errors(torch, c3["A"], dU, dH)
    assert bwd <= TOL and orth <= TOL, (prec, bwd, orth)
    n = dU.shape[0]
    d64 = (torch.linalg.norm(dU - c3["U"]) / math.sqrt(n)).item()
    assert d64 <= 10 * max(c3["d_cf"], 1e-13), (prec, d64, c3["d_cf"])
    dh = (torch.linalg.norm(dH - c3["H"]) / torch.linalg.norm(c3["H"])).item()
    assert dh <= 1e-13, (prec, dh)


def test_zolo_svd_n8192_sine(zl, torch_gpu):
    from zolo import workloads
    torch = torch_gpu
    n = 8192
    A, dA, U0, s, V0 = workloads.square(n, "sine", 0.0, 20260104)
    del U0, V0
    S, dU, dV, info, ph = zl.svd_device(dA)
    torch.cuda.synchronize()
    U, V = dU.T, dV.T
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    assert torch.max(torch.abs(S - torch.sort(s).values)).item() <= TOL
    assert (torch.linalg.norm(A - (U * S) @ V.T) / torch.linalg.norm(A)).item() <= TOL
    # the reference's residual ||A V - U diag(S)||_F / max(S) (check_routines.f90:70-86):
    # a Frobenius norm over n^2 entries that only divides by max(S), so it grows
    # as sqrt(n) times the relative error (measured 4.2e-13 at n = 8192)
    assert (torch.linalg.norm(A @ V - U * S) / S.max()).item() <= TOL * math.sqrt(n) / 10
    assert (torch.linalg.norm(U.T @ U - eye) / math.sqrt(n)).item() <= TOL
    assert (torch.linalg.norm(V.T @ V - eye) / math.sqrt(n)).item() <= TOL


@pytest.fixture(scope="module")
def c5(torch_gpu):
    """4096 FP32 matrices, n ~ U{32..512}, log10 kappa ~ U[1, 12]."""
    from zolo import workloads
    mats, ns, lk = workloads.batch(4096, 20260105)
    torch_gpu.cuda.synchronize()
    yield mats, ns
    del mats
    torch_gpu.cuda.empty_cache()


@pytest.mark.parametrize("l0", ["lu", "learned"])
def test_batched_4096(zl, torch_gpu, c5, l0):
    torch = torch_gpu
    mats, ns = c5
    if l0 == "learned":
        zl.l0_load_weights()
    Us, info = zl.pd_batched_device(mats, l0=l0)
    torch.cuda.synchronize()
    assert not info.failed.any()
    assert info.orth.max() <= 1e-6, info.orth.max()
    # the scalar plan of every matrix that neither fell back nor restarted
    plain = ~(info.fallback | info.continued)
    assert plain.mean() >= 0.95
    for k in np.flatnonzero(plain):
        _, sched = zo.batched_plan(info.beta[k])
        assert info.iters[k] == len(sched), (k, ns[k], info.beta[k], info.iters[k], sched)
    # independent FP64 check of a sample: orthogonality, backward error with
    # H = (U^T A + A^T U)/2, and H positive semidefinite
    rng = np.random.default_rng(5)
    for k in rng.choice(len(mats), 48, replace=False):
        A = mats[k].double().T
        U = Us[k].double().T
        n = A.shape[0]
        orth = (torch.linalg.norm(U.T @ U - torch.eye(n, dtype=torch.float64, device="cuda")) / math.sqrt(n)).item()
        assert orth <= 1e-6 and orth == pytest.approx(info.orth[k], rel=1e-2, abs=1e-9), (k, orth, info.orth[k])
        M = U.T @ A
        H = 0.5 * (M + M.T)
        nA = torch.linalg.norm(A).item()
        bwd = torch.linalg.norm(A - U @ H).item() / nA
        assert bwd <= 2e-5, (k, n, bwd)
        assert torch.linalg.eigvalsh(H).min().item() >= -2e-5 * nA, k
