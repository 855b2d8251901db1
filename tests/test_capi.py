"""CPU checks of the C ABI: the library loads, exports every symbol that
include/zolo.h declares, and its host-side scalar math (the Zolotarev
coefficients, the degree rule and the QDWH weight) equals the oracle's."""
import math
import os
import re

import numpy as np
import pytest

from oracle import zolo_oracle as zo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "zolo.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(zolo_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(zl):
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(zl.LIB, name), name


def test_version(zl):
    assert zl.LIB.zolo_version() >= 100


@pytest.mark.parametrize("con", [1.0005, 1.3, 1.9, 3.0, 150.0, 1e4, 1e6, 1e8, 1e12, 1e15])
@pytest.mark.parametrize("rk", [1, 2, 5, 8])
def test_coeff_matches_oracle(zl, con, rk):
    got = zl.compute_coeff(con, zo.DLAMCH_P, rk)
    ref = zo.compute_coeff(con, zo.DLAMCH_P, rk)
    assert np.allclose(got, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("con", [1.0, 1.001, 1.01, 1.1, 1.2, 1.5, 2.0, 6.49, 6.5, 179.9, 1.5e4, 2e6, 1e9, 3e12, 1e20])
def test_choosem_matches_oracle(zl, con):
    assert zl.choosem(con) == zo.choosem(con)


@pytest.mark.parametrize("x", [1e-12, 1e-6, 0.01, 0.5, 0.999, 1.0])
def test_hfunc_matches_oracle(zl, x):
    assert math.isclose(zl.hfunc(x), zo.hfunc(x), rel_tol=1e-14)


def test_elliptic_matches_oracle(zl):
    for alpha in [0.1, 1.0, 1.5, math.acos(1e-8)]:
        k, e = zl.mellipke(alpha, zo.DLAMCH_P)
        ko, eo = zo.mellipke(alpha, zo.DLAMCH_P)
        assert math.isclose(k, ko, rel_tol=1e-14) and math.isclose(e, eo, rel_tol=1e-12, abs_tol=1e-14)
        for u in [0.1, 0.7 * k]:
            got = zl.mellipj(u, alpha, zo.DLAMCH_P)
            ref = zo.mellipj(u, alpha, zo.DLAMCH_P)
            assert np.allclose(got, ref, rtol=1e-13, atol=1e-15)


def test_argument_validation_without_gpu(zl):
    with pytest.raises(zl.ZoloError) as e:
        zl.compute_coeff(0.5, 1e-16, 3)
    assert e.value.status == -1
    with pytest.raises(zl.ZoloError):
        zl.compute_coeff(2.0, 1e-16, 9)


def test_spectral_division_argument_codes_without_gpu(zl):
    """The spectral-division entry points check their arguments before any HIP call."""
    import ctypes as C
    L = zl.LIB
    buf = (C.c_double * 16)()
    rk, it = C.c_int(0), C.c_int(0)
    err = C.c_double(0.0)
    cnt = C.c_int(0)
    p = C.cast(buf, C.c_void_p)
    # zolo_subit_device(n, A, lda, normA, Q, ldq, rk, err, iters, stream)
    assert L.zolo_subit_device(0, p, 1, 1.0, p, 1, C.byref(rk), C.byref(err), C.byref(it), None) == -1
    assert L.zolo_subit_device(4, None, 4, 1.0, p, 4, C.byref(rk), C.byref(err), C.byref(it), None) == -2
    assert L.zolo_subit_device(4, p, 3, 1.0, p, 4, C.byref(rk), C.byref(err), C.byref(it), None) == -3
    assert L.zolo_subit_device(4, p, 4, 0.0, p, 4, C.byref(rk), C.byref(err), C.byref(it), None) == -4
    assert L.zolo_subit_device(4, p, 4, 1.0, p, 3, C.byref(rk), C.byref(err), C.byref(it), None) == -6
    assert L.zolo_subit_device(4, p, 4, 1.0, p, 4, None, C.byref(err), C.byref(it), None) == -7
    # zolo_eig_device(n, Q, ldq, As, lda, w, V, ldv, rk, err, iters, stream)
    assert L.zolo_eig_device(4, p, 4, p, 4, p, p, 3, C.byref(rk), None, None, None) == -8
    assert L.zolo_eig_device(4, p, 4, p, 4, None, p, 4, C.byref(rk), None, None, None) == -6
    # zolo_eig_below_device(precision, rk, n, A, lda, sigma, w, V, ldv, count, info, err, iters, stream)
    assert L.zolo_eig_below_device(5, 0, 4, p, 4, 0.0, p, p, 4, C.byref(cnt), None, None, None, None) == -1
    assert L.zolo_eig_below_device(0, 9, 4, p, 4, 0.0, p, p, 4, C.byref(cnt), None, None, None, None) == -2
    assert L.zolo_eig_below_device(0, 0, 0, p, 4, 0.0, p, p, 4, C.byref(cnt), None, None, None, None) == -3
    assert L.zolo_eig_below_device(0, 0, 4, p, 4, float("nan"), p, p, 4, C.byref(cnt), None, None, None,
                                   None) == -6
    assert L.zolo_eig_below_device(0, 0, 4, p, 4, 0.0, p, p, 4, None, None, None, None, None) == -10
    # zolo_eig_below_ex_device(precision, bounds, rk, n, ...): one position later, bounds is -2
    assert L.zolo_eig_below_ex_device(0, 2, 0, 4, p, 4, 0.0, p, p, 4, C.byref(cnt), None, None, None, None) == -2
    assert L.zolo_eig_below_ex_device(0, 1, 9, 4, p, 4, 0.0, p, p, 4, C.byref(cnt), None, None, None, None) == -3
    assert L.zolo_eig_below_ex_device(0, 1, 0, 4, p, 4, 0.0, p, p, 4, None, None, None, None, None) == -11
    # zolo_svd_device(precision, rk, n, A, lda, S, U, ldu, V, ldv, info, phase_ms, stream) and the _ex form
    assert L.zolo_svd_device(0, 0, 4, p, 4, p, p, 4, p, 3, None, None, None) == -10
    assert L.zolo_svd_ex_device(0, 5, 0, 4, p, 4, p, p, 4, p, 4, None, None, None) == -2
    assert L.zolo_svd_ex_device(0, 1, 0, 4, p, 4, p, p, 4, p, 3, None, None, None) == -11
    assert L.zolo_svd_ex_device(0, 0, 0, 4, None, 4, p, p, 4, p, 4, None, None, None) == -5
    # host forms
    d = C.cast(buf, C.POINTER(C.c_double))
    assert L.zolo_subit_host(4, d, 4, -1.0, d, 4, C.byref(rk), None, None) == -4
    assert L.zolo_eig_host(4, d, 4, d, 4, d, d, 4, None) == -9


def test_pd_argument_codes_without_gpu(zl):
    """The Zolo-PD and QDWH device entry points return their argument codes
    before any HIP call (a NULL A or U never reaches a kernel)."""
    import ctypes as C
    L = zl.LIB
    buf = (C.c_double * 16)()
    p = C.cast(buf, C.c_void_p)
    # zolo_pd_device(precision, sym, rk, n, alpha, beta, A, lda, U, ldu, H, ldh, info, stream)
    pd = L.zolo_pd_device
    assert pd(5, b"N", 2, 4, 1.0, 0.5, p, 4, p, 4, None, 4, None, None) == -1
    assert pd(0, b"X", 2, 4, 1.0, 0.5, p, 4, p, 4, None, 4, None, None) == -2
    assert pd(0, b"N", 9, 4, 1.0, 0.5, p, 4, p, 4, None, 4, None, None) == -3
    assert pd(0, b"N", 2, 0, 1.0, 0.5, p, 4, p, 4, None, 4, None, None) == -4
    assert pd(0, b"N", 2, 4, 0.0, 0.5, p, 4, p, 4, None, 4, None, None) == -5
    assert pd(0, b"N", 2, 4, 1.0, 2.0, p, 4, p, 4, None, 4, None, None) == -6
    assert pd(0, b"N", 2, 4, 1.0, 0.5, None, 4, p, 4, None, 4, None, None) == -7
    assert pd(0, b"N", 2, 4, 1.0, 0.5, p, 3, p, 4, None, 4, None, None) == -8
    assert pd(0, b"N", 2, 4, 1.0, 0.5, p, 4, None, 4, None, 4, None, None) == -9
    assert pd(0, b"N", 2, 4, 1.0, 0.5, p, 4, p, 3, None, 4, None, None) == -10
    assert pd(0, b"N", 2, 4, 1.0, 0.5, p, 4, p, 4, p, 3, None, None) == -12
    # zolo_pd_auto_device(precision, sym, rk, n, A, lda, U, ldu, H, ldh, info, stream): its own positions
    au = L.zolo_pd_auto_device
    assert au(5, b"N", 0, 4, p, 4, p, 4, None, 4, None, None) == -1
    assert au(0, b"N", -1, 4, p, 4, p, 4, None, 4, None, None) == -3
    assert au(0, b"N", 0, 0, p, 4, p, 4, None, 4, None, None) == -4
    assert au(0, b"N", 0, 4, None, 4, p, 4, None, 4, None, None) == -5
    assert au(0, b"N", 0, 4, p, 3, p, 4, None, 4, None, None) == -6
    assert au(0, b"N", 0, 4, p, 4, None, 4, None, 4, None, None) == -7
    assert au(0, b"N", 0, 4, p, 4, p, 3, None, 4, None, None) == -8
    assert au(0, b"N", 0, 4, p, 4, p, 4, p, 3, None, None) == -10
    # zolo_qdwh_device(sym, n, sigma, A, lda, U, ldu, H, ldh, info, stream)
    qd = L.zolo_qdwh_device
    assert qd(b"N", 4, 0.0, None, 4, p, 4, None, 4, None, None) == -4
    assert qd(b"N", 4, 0.0, p, 4, None, 4, None, 4, None, None) == -6
    assert qd(b"N", 4, float("inf"), p, 4, p, 4, None, 4, None, None) == -3


def test_mm_size_line_beyond_the_file_is_refused(zl, tmp_path):
    """A size line that claims more entries than the file can hold is an
    argument error (-3), not an allocation of that many entries."""
    f = tmp_path / "huge.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n3 3 9000000000000000000\n1 1 2.0\n")
    with pytest.raises(zl.ZoloError) as e:
        zl.mm_read(str(f))
    assert e.value.status == -3
    g = tmp_path / "short.mtx"
    g.write_text("%%MatrixMarket matrix coordinate real general\n3 3 5\n1 1 2.0\n2 2 3.0\n")
    with pytest.raises(zl.ZoloError) as e:
        zl.mm_read(str(g))
    assert e.value.status in (-3, -4)
