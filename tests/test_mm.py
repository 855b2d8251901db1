"""MatrixMarket input (prepare_matrix.f90:154-321, :638-759, mmread :1076):
the oracle's assembly against scipy.io.mmread, the library's host reader
against the oracle bit for bit, and the reader's error codes.  CPU only."""
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from oracle import zolo_oracle as zo


def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


def _random_sym(n, density, seed):
    M = sp.random(n, n, density=density, random_state=seed, format="coo")
    return (M + M.T + sp.eye(n)).tocoo()


@pytest.mark.parametrize("kind", ["symmetric", "general", "integer", "pattern"])
def test_oracle_vs_scipy(tmp_path, kind):
    p = os.path.join(str(tmp_path), f"{kind}.mtx")
    if kind == "general":
        M = sp.random(40, 40, density=0.15, random_state=3)
        sio.mmwrite(p, M)
    elif kind == "integer":
        M = _random_sym(30, 0.1, 4)
        M.data = np.round(M.data * 10)
        sio.mmwrite(p, M.astype(np.int64), symmetry="symmetric")
    elif kind == "pattern":
        M = _random_sym(30, 0.1, 5)
        sio.mmwrite(p, M, field="pattern", symmetry="symmetric")
    else:
        sio.mmwrite(p, _random_sym(60, 0.08, 6), symmetry="symmetric")
    sigma = 0.375
    A, mtype = zo.mm_dense(p, sigma)
    ref = sio.mmread(p).toarray().astype(np.float64)
    if mtype == 0:
        ref[np.diag_indices(ref.shape[0])] -= sigma
    assert mtype == (1 if kind == "general" else 0)
    assert np.array_equal(A, ref)


def test_oracle_mutation_diagonal_not_halved(tmp_path):
    p = os.path.join(str(tmp_path), "s.mtx")
    sio.mmwrite(p, _random_sym(20, 0.2, 7), symmetry="symmetric")
    A, _ = zo.mm_dense(p, 0.0, halve_diagonal=False)
    assert not np.array_equal(A, sio.mmread(p).toarray())


@pytest.mark.parametrize("kind", ["symmetric", "general", "pattern"])
def test_library_host_matches_oracle(zl, tmp_path, kind):
    p = os.path.join(str(tmp_path), "m.mtx")
    if kind == "general":
        sio.mmwrite(p, sp.random(70, 70, density=0.05, random_state=8))
    elif kind == "pattern":
        sio.mmwrite(p, _random_sym(50, 0.05, 9), field="pattern", symmetry="symmetric")
    else:
        sio.mmwrite(p, _random_sym(90, 0.05, 10), symmetry="symmetric")
    A, t = zl.mm_read(p, -1.25)
    R, tr = zo.mm_dense(p, -1.25)
    assert t == tr and np.array_equal(A, R)
    info = zl.mm_info(p, bandwidth=True)
    rows, cols = np.nonzero(sio.mmread(p).toarray() if kind == "general" else np.tril(sio.mmread(p).toarray()))
    assert info["rows"] == A.shape[0] and info["bandwidth"] == np.abs(rows - cols).max() + 1


def test_library_repeats_comments_and_d_exponents(zl, tmp_path):
    text = ("%%MatrixMarket MATRIX Coordinate REAL Symmetric\n% a comment\n\n% another\n"
            "4 4 6\n1 1 2.5D+00\n2 1 -1.0\n\n3 2 7.25e-1\n% inside\n2 1 3.0d0\n4 4 1\n4 3 0.5\n")
    p = _write(tmp_path, "r.mtx", text)
    A, t = zl.mm_read(p, 0.5)
    R, _ = zo.mm_dense(p, 0.5)
    assert t == 0 and np.array_equal(A, R)
    assert A[1, 0] == 3.0 and A[0, 1] == 3.0          # the later of the two (2,1) entries
    assert A[0, 0] == 2.0 and A[3, 3] == 0.5          # stored value - sigma


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1 0\n", -2),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -2),
    ("%%MatrixMarket matrix coordinate real unknown\n2 2 1\n1 1 1\n", -2),
    ("MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1\n", -2),
    ("%%MatrixMarket matrix coordinate real general\n2 x 1\n1 1 1\n", -3),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 2 1\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 q 1\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", -5),
    ("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1\n", -6),
])
def test_library_errors(zl, tmp_path, text, code):
    p = _write(tmp_path, "e.mtx", text)
    with pytest.raises(zl.ZoloError) as ei:
        zl.mm_read(p)
    assert ei.value.status == code


def test_library_missing_file(zl, tmp_path):
    with pytest.raises(zl.ZoloError) as ei:
        zl.mm_info(os.path.join(str(tmp_path), "absent.mtx"))
    assert ei.value.status == -1
