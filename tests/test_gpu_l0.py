"""Learned lower bound: device features, network and training against the CPU oracle.

The forward, backward and Adam kernels are checked against oracle/l0_oracle.py
(FP64 numpy, itself pinned by finite differences); the trained network is
compared with the LU-based estimate it replaces on the batched workload.
"""
import math
import os

import numpy as np
import pytest

from oracle import l0_oracle as lo
from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu


def _mats(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(sizes):
        kappa = 10.0 ** (1 + 2 * (i % 5))
        U = zo.haar(n, rng)
        V = zo.haar(n, rng)
        s = np.geomspace(1.0, 1.0 / kappa, n) if n > 1 else np.ones(1)
        out.append(((U * s) @ V.T).astype(np.float32))
    return out


def test_features_match_oracle(zl, torch_gpu):
    mats = _mats([1, 5, 8, 40, 100, 257, 512], 21)
    feat, smax = zl.l0_features_device([zl.to_device(A) for A in mats])
    feat = feat.cpu().numpy()
    smax = smax.cpu().numpy()
    for k, A in enumerate(mats):
        f, sm = lo.features(A.astype(np.float64))
        assert np.allclose(feat[k], f, rtol=1e-5, atol=2e-4), (A.shape[0], feat[k] - f)
        assert smax[k] == pytest.approx(sm, rel=1e-5)


def _params(seed):
    rng = np.random.default_rng(seed)
    return lo.init_params(rng, mu=rng.standard_normal(16) * 0.2, sd=0.5 + rng.random(16))


def test_forward_and_gradient_match_oracle(zl, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(5)
    P = _params(6)
    F = rng.standard_normal((301, 16))
    y0, _ = lo.forward(P, F)
    # targets away from the prediction, so float32 rounding cannot flip a pinball branch
    t = y0 + np.where(rng.random(301) < 0.5, -1.0, 1.0) * (0.5 + rng.random(301))
    dP = torch.tensor(P, dtype=torch.float32, device="cuda")
    dF = torch.tensor(F, dtype=torch.float32, device="cuda")
    dT = torch.tensor(t, dtype=torch.float32, device="cuda")
    y = zl.l0_forward_device(dP, dF).cpu().numpy()
    assert np.allclose(y, y0, rtol=1e-4, atol=1e-4)
    for tau in (0.05, 0.5):
        g, loss = zl.l0_grad_device(dP, dF, dT, tau)
        lref, gref = lo.loss_grad(P.astype(np.float32).astype(np.float64), F.astype(np.float32).astype(np.float64),
                                  t.astype(np.float32).astype(np.float64), tau)
        g = g.cpu().numpy()
        assert loss == pytest.approx(lref, rel=1e-4)
        assert np.max(np.abs(g - gref)) <= 1e-4 * np.max(np.abs(gref)), tau


def test_adam_step_and_training_progress(zl, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(8)
    P = _params(9)
    F = rng.standard_normal((512, 16))
    t = np.tanh(F[:, 0]) - 0.5 * F[:, 1] ** 2 + 0.1 * rng.standard_normal(512)
    dF = torch.tensor(F, dtype=torch.float32, device="cuda")
    dT = torch.tensor(t, dtype=torch.float32, device="cuda")
    dP = torch.tensor(P, dtype=torch.float32, device="cuda")
    h1 = zl.l0_train_device(dP, dF, dT, 1, lr=1e-3, tau=0.2)
    P32 = P.astype(np.float32).astype(np.float64)
    _, g = lo.loss_grad(P32, F.astype(np.float32).astype(np.float64), t.astype(np.float32).astype(np.float64), 0.2)
    Pref, _, _ = lo.adam_step(P32, g, np.zeros_like(g), np.zeros_like(g), 1e-3, 1)
    assert np.allclose(dP.cpu().numpy(), Pref, rtol=0, atol=2e-6)
    assert h1[0] == pytest.approx(lo.pinball(t, lo.forward(P32, F)[0], 0.2), rel=1e-4)
    dP = torch.tensor(P, dtype=torch.float32, device="cuda")
    hist = zl.l0_train_device(dP, dF, dT, 400, lr=3e-3, tau=0.2)
    assert hist[-1] < 0.5 * hist[0]


WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "zolo-svd_amd", "zolo",
                       "l0_weights.json")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained l0 weights file")
def test_learned_bound_gain_over_lu_estimate(zl, torch_gpu):
    """What the trained network gains on the batched workload (seeded batch of 512).

    The learned bound replaces an FP64 LU + condition estimate per matrix with
    cheap features and a network: its bound phase must be faster, its bounds
    tighter (a smaller CON needs fewer or cheaper steps) and the fallback to the
    LU estimate rare, with every matrix meeting the orthogonality target.
    """
    from zolo import workloads
    meta = zl.l0_load_weights(WEIGHTS)
    mats, ns, lk = workloads.batch(512, 777, nmax=384)
    _, info_lu = zl.pd_batched_device(mats, l0="lu")
    _, info_l0 = zl.pd_batched_device(mats, l0="learned")       # warm (graph / tile caches)
    _, info_lu = zl.pd_batched_device(mats, l0="lu")
    _, info_l0 = zl.pd_batched_device(mats, l0="learned")
    assert not info_lu.failed.any() and not info_l0.failed.any()
    assert info_l0.orth.max() <= 1e-6 and info_lu.orth.max() <= 1e-6
    fb = info_l0.fallback.mean()
    assert fb <= 0.05, fb
    t_lu = info_lu.phase_ms["bounds"]
    t_l0 = info_l0.phase_ms["learned"] + info_l0.phase_ms["bounds"]
    assert t_l0 < t_lu, (t_l0, t_lu)
    tight = np.median(np.log10(info_l0.beta / info_lu.beta))
    assert tight > 0.0, tight
    print(f"learned l0: bounds {t_l0:.1f} ms vs LU {t_lu:.1f} ms; total {info_l0.time_ms:.1f} vs {info_lu.time_ms:.1f} ms;"
          f" fallback {fb:.3f}; median tightening {tight:.2f} decades; mean iters {info_l0.iters.mean():.2f} vs "
          f"{info_lu.iters.mean():.2f}; margin {meta['margin']:.2f}")
