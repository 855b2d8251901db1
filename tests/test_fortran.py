"""The Fortran drop-in layer: libzolo_f.so exports the reference's entry
points (F77 names, trailing underscore) and module auxil_zolo; on the GPU the
Fortran driver (shaped like test_zolodrv1.f90 / pqdwh_fall.f90) calls
PZOLOPD1 and PQDWHFACS through the ISO_C_BINDING layer."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "zolo-svd_amd")

ENTRY = ["pzolopd1_", "pzolopd2_", "pqdwhfacs_", "zolochol_", "zoloqr_", "pdnormest_", "mpdgeqrf_", "mpdorgqr_",
         "formx2_", "hfunc_", "elpa_pdsyevd2_", "pqdwhsubit_", "zoloeig_"]


def test_fortran_library_exports_entry_points(zl):
    # zl first: torch's HIP runtime must be the one libzolo.so binds to
    lib = ctypes.CDLL(os.path.join(PKG, "libzolo_f.so"))
    for name in ENTRY:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", os.path.join(PKG, "libzolo_f.so")], capture_output=True, text=True).stdout
    for name in ["auxil_zolo", "compute_coeff", "choosem", "mellipke", "mellipj", "mod_prepare_matrix",
                 "prepare_real_symm_mm", "prepare_real_gem"]:
        assert name in out


def test_fortran_hfunc_on_host(zl):
    lib = ctypes.CDLL(os.path.join(PKG, "libzolo_f.so"))
    lib.hfunc_.restype = ctypes.c_double
    x = ctypes.c_double(1.0)
    assert abs(lib.hfunc_(ctypes.byref(x)) - 3.0) < 1e-15


@pytest.mark.gpu
def test_fortran_driver_on_gpu():
    exe = os.path.join(PKG, "test_pzolopd1")
    assert os.path.exists(exe), "build the Fortran driver first (make -C zolo-svd_amd)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout.splitlines()[-1]


@pytest.mark.gpu
def test_fortran_dropins_on_gpu():
    """ZOLOCHOL, ZOLOQR, PDNORMEST, MPDGEQRF / MPDORGQR, FormX2, PZOLOPD2 and
    PZOLOPD1 through their Fortran argument lists, submatrices at IA / JA > 1."""
    exe = os.path.join(PKG, "test_dropins")
    assert os.path.exists(exe), "build the Fortran driver first (make -C zolo-svd_amd)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout.splitlines()[-1]


@pytest.mark.gpu
def test_fortran_spectral_division_driver_on_gpu(tmp_path):
    """pqdwheig.f90's flow through the Fortran drop-ins: prepare_real_symm_MM,
    PQDWHFACS with a shift, PQDWHSUBIT, ELPA_PDSYEVD2 on the Rk block, ZoloEig."""
    exe = os.path.join(PKG, "test_zoloeig")
    assert os.path.exists(exe), "build the Fortran driver first (make -C zolo-svd_amd)"
    r = subprocess.run([exe, os.path.join(str(tmp_path), "a.mtx")], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout.splitlines()[-1]


def test_fortran_spectral_division_info_codes(zl):
    """PQDWHSUBIT and ZoloEig check their arguments before any GPU work."""
    lib = ctypes.CDLL(os.path.join(PKG, "libzolo_f.so"))
    I = ctypes.c_int
    D = ctypes.c_double
    buf = (D * 64)()
    desc = (I * 9)(1, 0, 4, 4, 4, 4, 0, 0, 4)
    bad = (I * 9)(1, 0, 4, 4, 4, 4, 1, 0, 4)   # RSRC /= 0
    info, rk = I(99), I(99)

    def subit(n, da, dq, norma):
        lib.pqdwhsubit_(ctypes.byref(I(n)), buf, ctypes.byref(I(1)), ctypes.byref(I(1)), da, ctypes.byref(D(norma)),
                        buf, ctypes.byref(I(1)), ctypes.byref(I(1)), dq, buf, ctypes.byref(I(16)), buf, buf,
                        ctypes.byref(I(16)), ctypes.byref(rk), ctypes.byref(info))
        return info.value

    assert subit(-1, desc, desc, 1.0) == -1
    assert subit(4, bad, desc, 1.0) == -507
    assert subit(4, desc, bad, 1.0) == -1007
    assert subit(4, desc, desc, 0.0) == -6
    assert subit(0, desc, desc, 1.0) == 0 and rk.value == 0

    def eig(n, lwork):
        lib.zoloeig_(ctypes.byref(I(n)), ctypes.byref(I(4)), buf, buf, ctypes.byref(I(0)), ctypes.byref(I(0)), buf,
                     ctypes.byref(I(lwork)), buf, ctypes.byref(I(16)), buf, ctypes.byref(info))
        return info.value

    assert eig(-2, 64) == -1
    assert eig(4, 15) == -8
    assert eig(0, 0) == 0
