"""Batched small-matrix Zolo-PD (FP32 data, one captured hipGraph) against the oracle.

Each matrix is compared with the FP64 CPU oracle on the same FP32 data: the
lower bound beta (LU + 1-norm estimate, test_zolodrv1.f90:179-189), the
planned schedule, orthogonality, backward error and, where the polar factor is
well conditioned, the distance to the SVD polar factor.
"""
import math

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu

# FP32 iterations: orthogonality target of the batched path and the backward
# error an FP32 polar decomposition reaches (a few unit roundoffs times sqrt(n)).
ORTH_TOL = 1e-6
BWD_TOL = 2e-5
SIZES = [1, 2, 7, 31, 32, 33, 63, 64, 65, 96, 127, 128, 129, 200, 255, 256, 257, 300, 384, 511, 512]
LOGK = [1.0, 3.0, 5.0, 6.5, 9.0, 11.0]


def _batch(seed=7, sizes=SIZES):
    rng = np.random.default_rng(seed)
    mats = []
    for i, n in enumerate(sizes):
        kappa = 10.0 ** LOGK[i % len(LOGK)]
        U = zo.haar(n, rng)
        V = zo.haar(n, rng)
        s = np.geomspace(1.0, 1.0 / kappa, n) if n > 1 else np.ones(1)
        mats.append(((U * s) @ V.T).astype(np.float32))
    return mats


def _run(zl, mats, **kw):
    import torch
    dA = [zl.to_device(A) for A in mats]
    dU, info = zl.pd_batched_device(dA, **kw)
    torch.cuda.synchronize()
    return [zl.from_device(u) for u in dU], info, dA, dU


def _check_polar(A32, U32, tol_orth=ORTH_TOL, tol_bwd=BWD_TOL):
    A = A32.astype(np.float64)
    U = U32.astype(np.float64)
    n = A.shape[0]
    H = 0.5 * (U.T @ A + A.T @ U)
    bwd = np.linalg.norm(A - U @ H) / np.linalg.norm(A)
    orth = np.linalg.norm(U.T @ U - np.eye(n)) / math.sqrt(n)
    assert orth <= tol_orth, (n, orth)
    assert bwd <= tol_bwd, (n, bwd)
    return orth, bwd


def test_batched_lu_bounds_schedule_and_polar(zl, torch_gpu):
    mats = _batch()
    Us, info, _, _ = _run(zl, mats, precision="f32", l0="lu")
    assert not info.failed.any() and not info.fallback.any()
    assert info.graph_nodes > 0 and info.buckets == len({max(32, -(-n // 32) * 32) for n in SIZES})
    for k, (A32, U32) in enumerate(zip(mats, Us)):
        n = A32.shape[0]
        _, be = zo.bounds_estimate(A32.astype(np.float64))
        assert info.beta[k] == pytest.approx(be, rel=1e-6), (n, info.beta[k], be)
        r, sched = zo.batched_plan(info.beta[k])
        assert info.iters[k] == len(sched), (n, info.iters[k], sched)
        orth, _ = _check_polar(A32, U32)
        assert info.orth[k] == pytest.approx(orth, rel=1e-3, abs=1e-12)
        kappa = 10.0 ** LOGK[k % len(LOGK)]
        if kappa <= 1e3 and n > 1:
            W, _, Vt = np.linalg.svd(A32.astype(np.float64))
            dist = np.linalg.norm(U32 - W @ Vt) / math.sqrt(n)
            assert dist <= 100 * 6e-8 * kappa, (n, kappa, dist)


def test_batched_graph_matches_eager(zl, torch_gpu):
    mats = _batch(seed=11, sizes=[40, 96, 130, 257])
    Ug, ig, _, _ = _run(zl, mats, use_graph=True)
    Ue, ie, _, _ = _run(zl, mats, use_graph=False)
    assert ie.graph_nodes == 0
    for a, b in zip(Ug, Ue):
        assert np.array_equal(a, b)
    assert np.array_equal(ig.iters, ie.iters)


def test_batched_fixed_tiles_agree(zl, torch_gpu):
    mats = _batch(seed=5, sizes=[130, 200])
    res = []
    for tile in (0, 1, 4):
        Us, info, _, _ = _run(zl, mats, tile=tile)
        assert not info.failed.any()
        for A32, U32 in zip(mats, Us):
            _check_polar(A32, U32)
        res.append(Us)
    assert zl.batched_tile(130) in (-1, 0, 1, 4)


def test_batched_fallback_on_overestimated_bound(zl, torch_gpu):
    mats = _batch(seed=13, sizes=[48, 100, 160, 224, 320, 64])
    _, info0, _, _ = _run(zl, mats)
    beta = info0.beta.copy()
    bad = np.arange(len(mats)) % 2 == 0
    beta_given = np.where(bad, np.minimum(beta * 1e4, 1.0), beta)
    Us, info, _, _ = _run(zl, mats, l0="given", beta=beta_given)
    # an overestimated sigma_min leaves singular values outside the planned
    # interval: the orthogonality check fails and the LU estimate takes over,
    # from the iterate (same polar factor) when the given plan ran at least one
    # step, else (beta_given = 1: nothing to iterate) from A
    assert np.array_equal(info.fallback, bad)
    assert not info.continued[~bad].any()
    assert info.continued[bad].sum() >= 2
    assert not info.failed.any()
    for k, (A32, U32) in enumerate(zip(mats, Us)):
        _check_polar(A32, U32)
        if bad[k] and not info.continued[k]:
            assert beta_given[k] == 1.0
            assert info.beta[k] == pytest.approx(beta[k], rel=1e-12)
        else:
            # the bound the plan started from is reported
            assert info.beta[k] == beta_given[k]


def test_batched_restart_near_singular(zl, torch_gpu):
    """A numerically singular FP32 matrix: CON ~ 6e15, where the plan is one
    FP64 QR step (the CON recurrence reports convergence after it) but that
    step leaves singular values in [0.12, 2.5] (the FP64 oracle shows the same);
    the matrix restarts from its iterate and meets the target."""
    rng = np.random.default_rng(23)
    n = 200
    A = (zo.haar(n, rng) * np.geomspace(1.0, 1e-3, n)) @ zo.haar(n, rng).T
    A[:, 1] = A[:, 0]
    A[7, 0] = 0.0
    A[7, 1] = 2e-13
    A32 = A.astype(np.float32)
    mats = [A32, _batch(seed=29, sizes=[64])[0]]
    Us, info, _, _ = _run(zl, mats)
    assert 1.0 / info.beta[0] > 1e15
    assert zo.batched_plan(info.beta[0])[1] == "Q"
    assert info.continued[0] and not info.continued[1]
    assert not info.failed.any()
    _check_polar(A32, Us[0])


def test_batched_argument_errors(zl, torch_gpu):
    import torch
    good = zl.to_device(np.eye(8, dtype=np.float32))
    big = torch.zeros((513, 513), dtype=torch.float32, device="cuda")
    with pytest.raises(zl.ZoloError) as e:
        zl.pd_batched_device([good, big])
    assert e.value.status == -2
    for prec in ("f64", "bf16x3"):
        with pytest.raises(zl.ZoloError) as e:
            zl.pd_batched_device([good], precision=prec)
        assert e.value.status == -7
    with pytest.raises(ValueError):
        zl.pd_batched_device([good], l0="given", beta=[0.1, 0.2])
