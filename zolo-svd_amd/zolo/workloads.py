"""Synthetic matrices of the workloads, generated on the GPU with torch.

``square``: one FP64 matrix A = U diag(s) V^T with Haar U, V and the spectra
of the full-size workloads (geometric, log-uniform, clustered, sine).

Batched polar decomposition workload: independent square FP32 matrices,
n drawn uniformly from {nmin..nmax}, A = U diag(s) V^T with U, V Haar
orthogonal (QR of a Gaussian matrix with the sign of diag(R) folded in),
log10(kappa) uniform in [lo, hi] and s geometric from 1 to 1/kappa.  The
matrices are formed in FP64 and rounded to FP32, and returned in the
column-major layout of ``zolo.to_device`` (row-major storage of A^T).
"""
from __future__ import annotations

import numpy as np


def batch(count: int, seed: int, nmin: int = 32, nmax: int = 512, logk=(1.0, 12.0)):
    """Returns (list of float32 CUDA tensors, n per matrix, log10 kappa per matrix)."""
    import torch
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count)
    lk = rng.uniform(logk[0], logk[1], size=count)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    mats = [None] * count
    groups = {}
    for i, n in enumerate(ns):
        groups.setdefault(int(n), []).append(i)
    for n in sorted(groups):
        idx = groups[n]
        G = torch.randn(2, len(idx), n, n, dtype=torch.float64, device="cuda", generator=g)
        Q, R = torch.linalg.qr(G)
        Q = Q * torch.sign(torch.diagonal(R, dim1=-2, dim2=-1)).unsqueeze(-2)
        kap = torch.tensor(lk[idx], dtype=torch.float64, device="cuda")
        t = torch.arange(n, dtype=torch.float64, device="cuda") / max(n - 1, 1)
        s = torch.pow(10.0, -kap[:, None] * t[None, :])
        A = (Q[0] * s[:, None, :]) @ Q[1].transpose(-1, -2).contiguous()
        At = A.transpose(-1, -2).contiguous().to(torch.float32)
        for j, i in enumerate(idx):
            mats[i] = At[j]
    return mats, ns, lk


def spectrum(kind: str, n: int, kappa: float, g):
    """Singular values (descending, FP64 CUDA tensor) of the square workloads.

    geometric: 1 -> 1/kappa; loguniform: s_1 = 1, s_n = 1/kappa pinned, the
    others 10^(-log10(kappa) u), u ~ U(0,1); cluster: geometric 1 -> 1/kappa
    over the first n - n//20 values, the last n//20 exactly 1/kappa; sine:
    1 + 0.5 sin(k) + 1e-3 N(0,1), k = 1..n.
    """
    import torch
    dev = "cuda"
    t = torch.arange(n, dtype=torch.float64, device=dev)
    if kind == "geometric":
        return torch.pow(kappa, -t / max(n - 1, 1))
    if kind == "loguniform":
        s = torch.pow(kappa, -torch.rand(n, dtype=torch.float64, device=dev, generator=g))
        s[0], s[-1] = 1.0, 1.0 / kappa
        return torch.sort(s, descending=True).values
    if kind == "cluster":
        m = n - n // 20
        s = torch.full((n,), 1.0 / kappa, dtype=torch.float64, device=dev)
        s[:m] = torch.pow(kappa, -t[:m] / max(m - 1, 1))
        return s
    if kind == "sine":
        return 1.0 + 0.5 * torch.sin(t + 1.0) + 1e-3 * torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    raise ValueError(kind)


def square(n: int, kind: str, kappa: float, seed: int):
    """A = U diag(s) V^T (FP64, on the GPU) with Haar U, V.

    Returns (A, dA, U, s, V): A row-major as torch holds it, dA the
    column-major buffer the library takes (the row-major buffer of A^T).
    """
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)

    def haar():
        Q, R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
        return Q * torch.sign(torch.diagonal(R))

    U = haar()
    V = haar()
    s = spectrum(kind, n, kappa, g)
    A = (U * s) @ V.T.contiguous()
    return A, A.T.contiguous(), U, s, V


def log_cond(mats, workers: int = 16) -> np.ndarray:
    """log10(sigma_min / sigma_max) of each FP32 matrix from an FP64 SVD (LAPACK on the host, threads)."""
    from concurrent.futures import ThreadPoolExecutor
    host = [m.double().cpu().numpy() for m in mats]

    def one(a):
        s = np.linalg.svd(a, compute_uv=False)
        return float(np.log10(s[-1] / s[0]))

    with ThreadPoolExecutor(max_workers=workers) as ex:
        return np.array(list(ex.map(one, host)))
