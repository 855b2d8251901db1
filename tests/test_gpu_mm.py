"""MatrixMarket input on the GPU: the device assembly (uploaded entries,
scatter, tiled A + A^T with the shifted diagonal) equals the host reader bit
for bit, and the reference's pqdwheig flow (read a symmetric sparse file,
eigenpairs below sigma) agrees with LAPACK dsyevd."""
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu


def _laplace2d(k):
    """5-point Laplacian on a k x k grid (a stiffness-like SPD matrix of order k^2)."""
    T = sp.diags([-np.ones(k - 1), 2 * np.ones(k), -np.ones(k - 1)], [-1, 0, 1])
    return (sp.kron(sp.eye(k), T) + sp.kron(T, sp.eye(k))).tocoo()


@pytest.mark.parametrize("n,kind", [(1, "symmetric"), (33, "symmetric"), (1000, "symmetric"), (257, "general"),
                                    (600, "repeats")])
def test_device_assembly_matches_host(zl, torch_gpu, tmp_path, n, kind):
    print(f"mm device assembly n={n} kind={kind}", flush=True)
    p = os.path.join(str(tmp_path), "m.mtx")
    if kind == "general":
        sio.mmwrite(p, sp.random(n, n, density=0.02, random_state=n))
    elif kind == "repeats":  # every entry written twice with different values: the later one counts
        M = sp.tril(sp.random(n, n, density=0.01, random_state=n) + sp.eye(n)).tocoo()
        lines = [f"{i + 1} {j + 1} {float(v)!r}" for i, j, v in zip(M.row, M.col, M.data)]
        lines += [f"{i + 1} {j + 1} {float(2.0 * v + 1.0)!r}" for i, j, v in zip(M.row, M.col, M.data)]
        with open(p, "w") as f:
            f.write(f"%%MatrixMarket matrix coordinate real symmetric\n{n} {n} {len(lines)}\n" + "\n".join(lines) + "\n")
    else:
        M = sp.tril(sp.random(n, n, density=0.02, random_state=n) + sp.eye(n)) if n > 1 else sp.coo_matrix([[3.5]])
        sio.mmwrite(p, M, symmetry="symmetric")
    sigma = 0.625
    A, t = zl.mm_read(p, sigma)
    dA, td = zl.mm_read_device(p, sigma)
    assert t == td
    assert np.array_equal(zl.from_device(dA), A)
    R, _ = zo.mm_dense(p, sigma)
    assert np.array_equal(A, R)


def test_eig_below_from_matrixmarket(zl, torch_gpu, tmp_path):
    """pqdwheig.f90 on one GPU: read the file, eigenpairs below sigma."""
    p = os.path.join(str(tmp_path), "lap.mtx")
    sio.mmwrite(p, sp.tril(_laplace2d(18)).tocoo(), symmetry="symmetric")
    dA, t = zl.mm_read_device(p)
    A = zl.from_device(dA)
    ev = np.linalg.eigvalsh(A)
    # sigma in the widest gap of the lower half of the spectrum
    gaps = np.diff(ev[: ev.size // 2])
    k = int(np.argmax(gaps))
    sigma = float(0.5 * (ev[k] + ev[k + 1]))
    w, V, count, _, err, iters = zl.eig_below_device(dA, sigma)
    assert t == 0 and count == k + 1
    w = w.cpu().numpy()
    V = zl.from_device(V)
    nrm = np.linalg.norm(A)
    assert np.max(np.abs(w - ev[: k + 1])) <= 1e-12 * nrm
    assert np.linalg.norm(A @ V - V * w) / nrm <= 1e-12
    assert np.linalg.norm(V.T @ V - np.eye(count)) <= 1e-12
