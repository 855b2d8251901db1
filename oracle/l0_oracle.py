"""CPU oracle of the learned lower bound (features, network, loss, Adam).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product path (the HIP library) never calls into it.

The learned part has no counterpart in the reference: it replaces the
condition estimate that sets l0 (test_zolodrv1.f90:185-189, PDGETRF +
PDGECON) for the batched path.  This module restates, in FP64 numpy, what the
library's kernels compute:

* ``features``  -- the 16 features of X = A / ||A||_F: log2 n, log10 of the 8
  Ritz values of X^T X after 8 Lanczos steps with full reorthogonalisation
  from a fixed +-1 start vector, log10 ||X||_1, log10 ||X||_inf, log10 |r_kk|
  of the R factor of X Omega (Omega: 4 columns of hash-seeded Box-Muller
  normals), and n * theta_1; plus smax = sqrt(theta_1).
* ``forward`` / ``loss_grad`` -- the 16 -> 64 -> 64 -> 1 tanh network and the
  gradient of the pinball loss mean rho_tau(t - y).
* ``adam_step`` -- one Adam update (beta1 0.9, beta2 0.999, eps 1e-8).

How each function is pinned (tests/test_l0_oracle.py): the gradient by
central finite differences; the pinball loss by its definition on hand-made
cases; the Lanczos Ritz values by eigvalsh of X^T X projected on the Krylov
basis and by Cauchy interlacing; the features by closed forms on matrices with
known spectra (a scaled orthogonal matrix, a diagonal matrix); Adam by its
first step, which moves every parameter by lr * sign(g).
"""
from __future__ import annotations

import math

import numpy as np

N_FEAT = 16
N_HID = 64
N_LANCZOS = 8
N_SKETCH = 4
N_TRAIN = N_HID * N_FEAT + N_HID + N_HID * N_HID + N_HID + N_HID + 1
N_PARAM = 2 * N_FEAT + N_TRAIN + 1
M32 = 0xFFFFFFFF


def hash32(x: np.ndarray) -> np.ndarray:
    """32-bit integer hash (xorshift-multiply), as the feature kernel."""
    x = np.asarray(x, dtype=np.uint64) & M32
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def _unit(key):
    return (hash32(key).astype(np.float64) + 0.5) * (1.0 / 4294967296.0)


def gauss(key: np.ndarray) -> np.ndarray:
    key = np.asarray(key, dtype=np.uint64)
    u1 = _unit(2 * key + 1)
    u2 = _unit(2 * key + 2)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def start_vector(n: int) -> np.ndarray:
    i = np.arange(n, dtype=np.uint64)
    s = np.where(hash32(i + 0x9E3779B9) & 1, 1.0, -1.0)
    return s / math.sqrt(n)


def sketch_matrix(n: int) -> np.ndarray:
    j = np.arange(n, dtype=np.uint64)[:, None]
    k = np.arange(N_SKETCH, dtype=np.uint64)[None, :]
    return gauss(j * N_SKETCH + k)


def lanczos(M: np.ndarray, q0: np.ndarray, k: int = N_LANCZOS):
    """k Lanczos steps on the symmetric M with full (twice) reorthogonalisation.

    Returns (alpha, beta, Q) with Q's columns the Lanczos vectors used.
    """
    n = M.shape[0]
    Q = [q0]
    a, b = [], []
    for j in range(min(k, n)):
        w = M @ Q[j]
        a.append(float(Q[j] @ w))
        for p in range(j + 1):
            for _ in range(2):
                w = w - (Q[p] @ w) * Q[p]
        bn = float(np.linalg.norm(w))
        b.append(bn)
        if bn <= 1e-14 * abs(a[-1]) or j + 1 == k:
            break
        Q.append(w / bn)
    return np.array(a), np.array(b), np.stack(Q, axis=1)


def ritz_values(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    k = len(a)
    T = np.diag(a) + np.diag(b[:k - 1], 1) + np.diag(b[:k - 1], -1)
    return np.sort(np.linalg.eigvalsh(T))[::-1]


def features(A: np.ndarray) -> tuple[np.ndarray, float]:
    """The 16 features and smax of one matrix (FP64 on the given data)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    X = A / np.linalg.norm(A)
    n1 = np.abs(X).sum(axis=0).max()
    ninf = np.abs(X).sum(axis=1).max()
    Y = X @ sketch_matrix(n)
    rd = np.zeros(N_SKETCH)
    for k in range(N_SKETCH):      # modified Gram-Schmidt
        for p in range(k):
            Y[:, k] -= (Y[:, p] @ Y[:, k]) * Y[:, p]
        rd[k] = np.linalg.norm(Y[:, k])
        Y[:, k] = Y[:, k] / rd[k] if rd[k] > 0 else 0.0
    a, b, _ = lanczos(X.T @ X, start_vector(n))
    th = ritz_values(a, b)
    th = np.concatenate([th, np.full(N_LANCZOS - len(th), th[-1])])
    f = np.zeros(N_FEAT)
    f[0] = math.log2(n)
    f[1:9] = np.log10(np.maximum(th, 1e-30))
    f[9] = math.log10(max(n1, 1e-30))
    f[10] = math.log10(max(ninf, 1e-30))
    f[11:15] = np.log10(np.maximum(rd, 1e-30))
    th1 = max(th[0], 1e-300)
    f[15] = n * th1
    return f, math.sqrt(th1)


# ---------------------------------------------------------------------------
# Network
# ---------------------------------------------------------------------------

def unpack(P: np.ndarray):
    P = np.asarray(P, dtype=np.float64)
    o = 0

    def take(k, shape=None):
        nonlocal o
        v = P[o:o + k]
        o += k
        return v.reshape(shape) if shape else v

    mu = take(N_FEAT)
    sd = take(N_FEAT)
    W1 = take(N_HID * N_FEAT, (N_HID, N_FEAT))
    b1 = take(N_HID)
    W2 = take(N_HID * N_HID, (N_HID, N_HID))
    b2 = take(N_HID)
    w3 = take(N_HID)
    b3 = take(1)[0]
    margin = take(1)[0]
    return mu, sd, W1, b1, W2, b2, w3, b3, margin


def forward(P, F):
    mu, sd, W1, b1, W2, b2, w3, b3, _ = unpack(P)
    Xn = (np.asarray(F, dtype=np.float64) - mu) / sd
    H1 = np.tanh(Xn @ W1.T + b1)
    H2 = np.tanh(H1 @ W2.T + b2)
    return H2 @ w3 + b3, (Xn, H1, H2)


def pinball(t, y, tau):
    r = np.asarray(t, dtype=np.float64) - np.asarray(y, dtype=np.float64)
    return float(np.mean(np.where(r >= 0, tau * r, (tau - 1.0) * r)))


def loss_grad(P, F, t, tau):
    """(loss, gradient w.r.t. the trained block W1, b1, W2, b2, w3, b3 in storage order)."""
    mu, sd, W1, b1, W2, b2, w3, b3, _ = unpack(P)
    y, (Xn, H1, H2) = forward(P, F)
    N = len(y)
    d3 = ((y > np.asarray(t)).astype(np.float64) - tau) / N
    d2 = np.outer(d3, w3) * (1.0 - H2 * H2)
    d1 = (d2 @ W2) * (1.0 - H1 * H1)
    g = np.concatenate([(d1.T @ Xn).ravel(), d1.sum(0), (d2.T @ H1).ravel(), d2.sum(0), H2.T @ d3, [d3.sum()]])
    return pinball(t, y, tau), g


def adam_step(P, g, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8):
    P = np.array(P, dtype=np.float64, copy=True)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    o = 2 * N_FEAT
    P[o:o + N_TRAIN] -= lr * mh / (np.sqrt(vh) + eps)
    return P, m, v


def beta_of(P, y, smax):
    margin = unpack(P)[-1]
    return 10.0 ** np.minimum(np.asarray(y) - margin, 0.0) * np.asarray(smax)


def init_params(rng: np.random.Generator, mu=None, sd=None):
    P = np.zeros(N_PARAM)
    P[:N_FEAT] = 0.0 if mu is None else mu
    P[N_FEAT:2 * N_FEAT] = 1.0 if sd is None else sd
    o = 2 * N_FEAT
    P[o:o + N_HID * N_FEAT] = rng.standard_normal(N_HID * N_FEAT) / math.sqrt(N_FEAT)
    o += N_HID * N_FEAT + N_HID
    P[o:o + N_HID * N_HID] = rng.standard_normal(N_HID * N_HID) / math.sqrt(N_HID)
    o += N_HID * N_HID + N_HID
    P[o:o + N_HID] = rng.standard_normal(N_HID) / math.sqrt(N_HID)
    return P
