"""The no-allocation PDNORMEST entry point (zolo_normest_noalloc_device)
against zolo_normest_device on the same inputs: the same kernels on a fixed
device array instead of the library workspace, so the estimate and the
iteration count agree bit for bit.  Two calls in a row, the second smaller,
share the one array."""
import ctypes as C

import numpy as np
import pytest

from oracle import zolo_oracle as zo

pytestmark = pytest.mark.gpu


def _noalloc(zl):
    f = zl.LIB.zolo_normest_noalloc_device
    f.restype = C.c_int
    f.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_double),
                  C.POINTER(C.c_int), C.c_void_p]
    return f


def _both(zl, torch, m, n, seed, tol=1e-6):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * np.geomspace(1.0, 1e-3, n)
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()   # column-major m x n, ld m
    e1, c1 = C.c_double(), C.c_int()
    st = zl.LIB.zolo_normest_device(m, n, C.c_void_p(dA.data_ptr()), m, tol, C.byref(e1), C.byref(c1), None)
    assert st == 0
    e2, c2 = C.c_double(), C.c_int()
    st = _noalloc(zl)(m, n, C.c_void_p(dA.data_ptr()), m, tol, C.byref(e2), C.byref(c2), None)
    assert st == 0, zl.LIB.zolo_last_error()
    torch.cuda.synchronize()
    assert e2.value == e1.value and c2.value == c1.value, (e1.value, e2.value, c1.value, c2.value)
    eo, co = zo.normest(A, tol)   # the power iteration of pdnormest.f90:43-80 on the host
    assert c2.value == co and abs(e2.value - eo) <= 1e-10 * eo, (e2.value, eo, c2.value, co)
    return e2.value


def test_noalloc_normest_matches_workspace_form(zl, torch_gpu):
    _both(zl, torch_gpu, 1500, 1200, 1)
    _both(zl, torch_gpu, 300, 77, 2)      # second call, smaller: the array keeps the first call's data


def test_noalloc_normest_rejects_oversized(zl, torch_gpu):
    dA = torch_gpu.zeros(8, dtype=torch_gpu.float64, device="cuda")
    e, c = C.c_double(), C.c_int()
    st = _noalloc(zl)(40000, 40000, C.c_void_p(dA.data_ptr()), 40000, 1e-6, C.byref(e), C.byref(c), None)
    assert st == -1
    assert b"more than the fixed" in zl.LIB.zolo_last_error()
