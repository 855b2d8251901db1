"""Python front end of the MI355X Zolo-PD library (ctypes over libzolo.so).

Mirrors the reference's Fortran entry points on a 1x1 process grid:

* ``pzolopd1``   -- PZOLOPD1 (pzolopd1.f90:2): Zolo-PD with caller-supplied alpha, beta.
* ``pqdwhfacs``  -- PQDWHFACS (pqdwhfacs.f90:2): QDWH-PD with its own bounds.
* ``zolochol``   -- ZOLOCHOL (zolochol.f90:2): one shifted Cholesky contribution.
* ``zoloqr``     -- ZOLOQR (zoloqr.f90:2): one structured-QR contribution.
* ``pdnormest``  -- PDNORMEST (pdnormest.f90:2): 2-norm power-iteration estimate.
* ``bounds``     -- the alpha/beta estimate of test_zolodrv1.f90:179-189.
* ``compute_coeff`` / ``choosem`` / ``hfunc`` / ``mellipke`` / ``mellipj`` -- Auxil_Zolo.f90, hfunc.f90.
* ``pd_batched_device`` -- PZOLOPD1 over many small FP32 matrices of uneven size in one hipGraph.
* ``syevd_device`` / ``svd_device`` -- the Zolo-SVD tail: ELPA_PDSYEVD2's role and ZoloCHK1's assembly.
* ``mm_info`` / ``mm_read`` / ``mm_read_device`` -- MatrixMarket input as prepare_real_GEM /
  prepare_real_symm_MM (prepare_matrix.f90:154, :638) assemble it.
* ``subit_device`` / ``eig_device`` / ``eig_below_device`` -- spectral division: PQDWHSUBIT
  (pqdwhsubit.f90:2), ZoloEig (ZoloEig.f90:2) and the driver of pqdwheig.f90.

Host arrays are numpy (column-major copies are made as needed); the
``*_device`` variants take torch CUDA tensors (column-major data in HBM, see
``to_device``) and run on torch's current stream.  Every call goes to the HIP
library; there is no CPU fallback: if libzolo.so is missing, importing this
module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(os.path.dirname(_HERE), "libzolo.so")

if not os.path.exists(_LIB_PATH):
    raise ImportError(f"libzolo.so not found at {_LIB_PATH}; build it with `make -C zolo-svd_amd`")
# One HIP runtime per process: torch ships its own libamdhip64 (found by torch
# under an unversioned name), so it has to be loaded first; libzolo.so then
# binds to that copy through the shared soname instead of loading a second
# runtime that would not see the device.
try:
    import torch as _torch_runtime  # noqa: F401
except ImportError:  # pragma: no cover - host-only use of the scalar helpers
    _torch_runtime = None
_lib = C.CDLL(_LIB_PATH)

PREC = {"f64": 0, "f32": 1, "bf16x3": 2, "f32_refsched": 3, "bf16x3_refsched": 4}
GEMM_PREC = {"f64": 0, "f32": 1, "bf16x3": 2, "f64x": 3}
ERR_NOT_SPD = 1002
ERR_NO_CONVERGENCE = 1003


class ZoloError(RuntimeError):
    def __init__(self, func: str, status: int):
        msg = _lib.zolo_last_error().decode(errors="replace")
        super().__init__(f"{func} failed with status {status}: {msg}")
        self.status = status


class _Info(C.Structure):
    _fields_ = [
        ("iters", C.c_int), ("r", C.c_int), ("info", C.c_int), ("polish_passes", C.c_int), ("restarts", C.c_int),
        ("schedule", C.c_char * 16),
        ("err", C.c_double * 16), ("con", C.c_double * 16), ("maxc", C.c_double * 16),
        ("alpha", C.c_double), ("beta", C.c_double), ("time_ms", C.c_double),
        ("phase_ms", C.c_double * 8), ("shifts", C.c_int * 16), ("cg_iters", C.c_int * 4),
    ]


PHASES = ("bounds", "qr", "chol", "formx2", "polish", "h", "other")


@dataclass
class RunInfo:
    """Iteration record of one call (zolo_info_t)."""
    iters: int
    r: int
    schedule: str
    err: list
    con: list
    maxc: list
    alpha: float
    beta: float
    time_ms: float
    phase_ms: dict = field(default_factory=dict)
    polish_passes: int = 0
    info: int = 0
    restarts: int = 0
    shifts: list = field(default_factory=list)
    cg_iters: list = field(default_factory=list)

    @staticmethod
    def from_c(ci: _Info) -> "RunInfo":
        k = ci.iters
        return RunInfo(iters=k, r=ci.r, schedule=ci.schedule.decode()[:k], err=list(ci.err[:k]),
                       con=list(ci.con[:k]), maxc=list(ci.maxc[:k]), alpha=ci.alpha, beta=ci.beta,
                       time_ms=ci.time_ms, phase_ms={p: ci.phase_ms[i] for i, p in enumerate(PHASES)},
                       polish_passes=ci.polish_passes, info=ci.info, restarts=ci.restarts,
                       shifts=list(ci.shifts[:k]), cg_iters=list(ci.cg_iters[:ci.polish_passes]))


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.c_int64


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("zolo_last_error", C.c_char_p, [])
_sig("zolo_init", C.c_int, [C.c_int])
_sig("zolo_finalize", C.c_int, [])
_sig("zolo_version", C.c_int, [])
_sig("zolo_compute_coeff", C.c_int, [C.c_double, C.c_double, C.c_int, _D])
_sig("zolo_choosem", C.c_int, [C.c_double])
_sig("zolo_hfunc", C.c_double, [C.c_double])
_sig("zolo_mellipke", None, [C.c_double, C.c_double, _D, _D])
_sig("zolo_mellipj", None, [C.c_double, C.c_double, C.c_double, _D, _D, _D])
_sig("zolo_pd_device", C.c_int, [C.c_int, C.c_char, C.c_int, _I64, C.c_double, C.c_double, _P, _I64, _P, _I64, _P,
                                 _I64, C.POINTER(_Info), _P])
_sig("zolo_pd_host", C.c_int, [C.c_int, C.c_char, C.c_int, _I64, C.c_double, C.c_double, _D, _I64, _D, _I64, _D,
                               _I64, C.POINTER(_Info)])
_sig("zolo_pd2_host", C.c_int, [C.c_int, C.c_char, C.c_int, _I64, C.c_double, C.c_double, _D, _I64, _D, _I64, _D,
                                _I64, C.POINTER(_Info)])
_sig("zolo_bounds_device", C.c_int, [C.c_int, _I64, _P, _I64, _D, _D, _P])
_sig("zolo_qdwh_device", C.c_int, [C.c_char, _I64, C.c_double, _P, _I64, _P, _I64, _P, _I64, C.POINTER(_Info), _P])
_sig("zolo_qdwh_host", C.c_int, [C.c_char, _I64, C.c_double, _D, _I64, _D, _I64, _D, _I64, C.POINTER(_Info)])
_sig("zolo_chol_step_host", C.c_int, [C.c_char, _I64, _D, _I64, _D, _I64, _D, C.c_int, C.c_int])
_sig("zolo_qr_step_host", C.c_int, [_I64, _D, _I64, _D, _I64, _D, C.c_int, C.c_int])
_sig("zolo_normest_device", C.c_int, [_I64, _I64, _P, _I64, C.c_double, _D, C.POINTER(C.c_int), _P])
_sig("zolo_gemm_device", C.c_int, [C.c_int, C.c_int, C.c_int, _I64, _I64, _I64, C.c_double, _P, _I64, _P, _I64,
                                   C.c_double, _P, _I64, _P])
_sig("zolo_stacked_qr_device", C.c_int, [_I64, _P, _I64, _P, _I64, _P, _P])
_sig("zolo_chol_inv_device", C.c_int, [C.c_int, _I64, _P, _I64, _P, _I64, _P])
_sig("zolo_pd_batched_device", C.c_int, [C.c_int, _P, _P, _P, _P, _P, C.c_int, C.c_int, _P, C.c_double, C.c_int,
                                         C.c_int, C.c_int, _P, _P, _P, _P, _P, _P])
_sig("zolo_batched_tile", C.c_int, [_I64, C.c_int])
_sig("zolo_syevd_device", C.c_int, [_I64, _P, _I64, _P, _P, _I64, _P, _P])
_sig("zolo_svd_device", C.c_int, [C.c_int, C.c_int, _I64, _P, _I64, _P, _P, _I64, _P, _I64, C.POINTER(_Info), _P,
                                  _P])
_sig("zolo_svd_ex_device", C.c_int, [C.c_int, C.c_int, C.c_int, _I64, _P, _I64, _P, _P, _I64, _P, _I64,
                                     C.POINTER(_Info), _P, _P])
_sig("zolo_pd_auto_device", C.c_int, [C.c_int, C.c_char, C.c_int, _I64, _P, _I64, _P, _I64, _P, _I64,
                                      C.POINTER(_Info), _P])
_sig("zolo_subit_device", C.c_int, [_I64, _P, _I64, C.c_double, _P, _I64, C.POINTER(C.c_int), _D,
                                    C.POINTER(C.c_int), _P])
_sig("zolo_eig_device", C.c_int, [_I64, _P, _I64, _P, _I64, _P, _P, _I64, C.POINTER(C.c_int), _D, C.POINTER(C.c_int),
                                  _P])
_sig("zolo_eig_below_device", C.c_int, [C.c_int, C.c_int, _I64, _P, _I64, C.c_double, _P, _P, _I64,
                                        C.POINTER(C.c_int), C.POINTER(_Info), _D, C.POINTER(C.c_int), _P])
_sig("zolo_eig_below_ex_device", C.c_int, [C.c_int, C.c_int, C.c_int, _I64, _P, _I64, C.c_double, _P, _P, _I64,
                                           C.POINTER(C.c_int), C.POINTER(_Info), _D, C.POINTER(C.c_int), _P])
_sig("zolo_subit_host", C.c_int, [_I64, _D, _I64, C.c_double, _D, _I64, C.POINTER(C.c_int), _D, C.POINTER(C.c_int)])
_sig("zolo_eig_host", C.c_int, [_I64, _D, _I64, _D, _I64, _D, _D, _I64, C.POINTER(C.c_int)])
_sig("zolo_mm_info", C.c_int, [C.c_char_p, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(C.c_int),
                               C.POINTER(_I64)])
_sig("zolo_mm_dense_host", C.c_int, [C.c_char_p, C.c_double, _D, _I64, C.POINTER(C.c_int)])
_sig("zolo_mm_dense_device", C.c_int, [C.c_char_p, C.c_double, _P, _I64, C.POINTER(C.c_int), _P])
_sig("zolo_l0_param_count", C.c_int, [])
_sig("zolo_l0_set_weights", C.c_int, [_P, C.c_int])
_sig("zolo_l0_features_device", C.c_int, [C.c_int, _P, _P, _P, _P, _P, _P])
_sig("zolo_l0_forward_device", C.c_int, [_P, _P, C.c_int, _P, _P])
_sig("zolo_l0_grad_device", C.c_int, [_P, _P, _P, C.c_int, C.c_float, _P, C.POINTER(C.c_float), _P])
_sig("zolo_l0_train_device", C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_float, C.c_float, _P, _P])
_sig("zolo_formx2_host", C.c_int, [_I64, C.c_double, _D, _I64, _D, _I64, _D])
_sig("zolo_profile_enable", C.c_int, [C.c_int])
_sig("zolo_set_f64_products", C.c_int, [C.c_int])
_sig("zolo_get_f64_products", C.c_int, [])
_sig("zolo_set_panel_exchange", C.c_int, [C.c_int])
_sig("zolo_get_panel_exchange", C.c_int, [])
_sig("zolo_profile_gemm", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64), _D, _D])


def set_f64_products(mode: int):
    """FP64 parts of the mixed precisions: 1 = exact int8 products (default), 0 = FP64 MFMA."""
    _check("zolo_set_f64_products", _lib.zolo_set_f64_products(int(mode)))


def get_f64_products() -> int:
    return int(_lib.zolo_get_f64_products())


def set_panel_exchange(mode: int):
    """QR panels of one shift per XCD: 0 = automatic (L2-local stores when the
    placement check passes), 1 = device-scope stores (the fallback form)."""
    _check("zolo_set_panel_exchange", _lib.zolo_set_panel_exchange(int(mode)))


def get_panel_exchange() -> int:
    return int(_lib.zolo_get_panel_exchange())


def profile_enable(on: bool = True):
    """Start (clearing) or stop the GEMM launch ledger."""
    _check("zolo_profile_enable", _lib.zolo_profile_enable(1 if on else 0))


def profile_gemm(precision: int = -1, transa: int = -1, transb: int = -1, tile_m: int = -1, tile_n: int = -1):
    """(launches, total ms, algorithmic flops) of the recorded GEMM launches matching the key (-1 = any)."""
    cnt, ms, fl = C.c_int64(), C.c_double(), C.c_double()
    _check("zolo_profile_gemm", _lib.zolo_profile_gemm(precision, transa, transb, tile_m, tile_n, C.byref(cnt),
                                                       C.byref(ms), C.byref(fl)))
    return cnt.value, ms.value, fl.value

LIB = _lib
LIB_PATH = _LIB_PATH


def _check(func: str, st: int):
    if st != 0:
        raise ZoloError(func, st)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _fortran(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------------------
# Scalar helpers (host code inside libzolo.so)
# ---------------------------------------------------------------------------

def compute_coeff(con: float, tol: float, rk: int) -> np.ndarray:
    """Compute_Coeff (Auxil_Zolo.f90:165): c_1..c_2r followed by a_1..a_r."""
    out = np.zeros(3 * rk)
    _check("zolo_compute_coeff", _lib.zolo_compute_coeff(con, tol, rk, _dptr(out)))
    return out


def choosem(con: float) -> int:
    """choosem (Auxil_Zolo.f90:130)."""
    return int(_lib.zolo_choosem(con))


def hfunc(x: float) -> float:
    """hfunc (hfunc.f90:2)."""
    return float(_lib.zolo_hfunc(x))


def mellipke(alpha: float, tol: float):
    k, e = C.c_double(), C.c_double()
    _lib.zolo_mellipke(alpha, tol, C.byref(k), C.byref(e))
    return k.value, e.value


def mellipj(u: float, alpha: float, tol: float):
    sn, cn, dn = C.c_double(), C.c_double(), C.c_double()
    _lib.zolo_mellipj(u, alpha, tol, C.byref(sn), C.byref(cn), C.byref(dn))
    return sn.value, cn.value, dn.value


# ---------------------------------------------------------------------------
# Host-array entry points (mirror the Fortran argument meaning)
# ---------------------------------------------------------------------------

def pzolopd1(A, alpha: float, beta: float, rk: int, sym: str = "N", polar: bool = True,
             precision: str = "f64"):
    """PZOLOPD1 on the GPU.  Returns (U, H or None, RunInfo).

    X0 = A/(alpha*beta), CON = 1/beta (pzolopd1.f90:286-291); beta must be a
    lower bound of sigma_min(A/alpha).  precision 'f64' is the reference's
    arithmetic; 'f32' / 'bf16x3' run the Cholesky iterations in low precision
    and finish with an FP64 polish.
    """
    A = _fortran(A)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("pzolopd1: A must be square (pzolopd1.f90:192)")
    U = np.zeros((n, n), order="F")
    H = np.zeros((n, n), order="F") if polar else None
    ci = _Info()
    st = _lib.zolo_pd_host(PREC[precision], sym.encode()[:1], rk, n, alpha, beta, _dptr(A), n, _dptr(U), n,
                           _dptr(H) if polar else None, n, C.byref(ci))
    _check("zolo_pd_host", st)
    return U, H, RunInfo.from_c(ci)


def pzolopd2(A, alpha: float, beta: float, rk: int, sym: str = "N", polar: bool = True,
             precision: str = "f64"):
    """PZOLOPD2 on the GPU: PZOLOPD1 with the iteration cap of 12 and the restart
    of pzolopd2.f90:397-405 (bounds of the iterate re-estimated when CON reaches
    1 before ERR meets TOL2).  Returns (U, H or None, RunInfo)."""
    A = _fortran(A)
    n = A.shape[0]
    U = np.zeros((n, n), order="F")
    H = np.zeros((n, n), order="F") if polar else None
    ci = _Info()
    st = _lib.zolo_pd2_host(PREC[precision], sym.encode()[:1], rk, n, alpha, beta, _dptr(A), n, _dptr(U), n,
                            _dptr(H) if polar else None, n, C.byref(ci))
    _check("zolo_pd2_host", st)
    return U, H, RunInfo.from_c(ci)


def pqdwhfacs(A, sigma: float = 0.0, polar: bool = True, sym: str = "N"):
    """PQDWHFACS on the GPU.  Returns (U, H or None, RunInfo)."""
    A = _fortran(A)
    n = A.shape[0]
    U = np.zeros((n, n), order="F")
    H = np.zeros((n, n), order="F") if polar else None
    ci = _Info()
    st = _lib.zolo_qdwh_host(sym.encode()[:1], n, sigma, _dptr(A), n, _dptr(U), n, _dptr(H) if polar else None, n,
                             C.byref(ci))
    _check("zolo_qdwh_host", st)
    return U, H, RunInfo.from_c(ci)


def zolochol(X, coeff, rk: int, ip: int, sym: str = "N") -> np.ndarray:
    """ZOLOCHOL: Z = -a_ip X (X^T X + c_{2ip-1} I)^{-1} (ip 1-based)."""
    X = _fortran(X)
    n = X.shape[0]
    Z = np.zeros((n, n), order="F")
    cf = np.ascontiguousarray(coeff, dtype=np.float64)
    _check("zolo_chol_step_host",
           _lib.zolo_chol_step_host(sym.encode()[:1], n, _dptr(X), n, _dptr(Z), n, _dptr(cf), rk, ip))
    return Z


def zoloqr(X, coeff, rk: int, ip: int) -> np.ndarray:
    """ZOLOQR: Z = -a_ip/sqrt(c) Q1 Q2^T with [X; sqrt(c) I] = QR (ip 1-based)."""
    X = _fortran(X)
    n = X.shape[0]
    Z = np.zeros((n, n), order="F")
    cf = np.ascontiguousarray(coeff, dtype=np.float64)
    _check("zolo_qr_step_host", _lib.zolo_qr_step_host(n, _dptr(X), n, _dptr(Z), n, _dptr(cf), rk, ip))
    return Z


# ---------------------------------------------------------------------------
# Device entry points (torch CUDA tensors holding column-major matrices)
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def to_device(A, dtype=None):
    """Column-major device copy: returns a torch tensor T with T.T laid out as A (ld = n)."""
    torch = _torch()
    a = np.asarray(A)
    dt = dtype or (torch.float64 if a.dtype == np.float64 else torch.float32)
    # A column-major n x m matrix is a row-major m x n array of its transpose.
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device="cuda", dtype=dt)


def from_device(t) -> np.ndarray:
    return t.detach().cpu().numpy().T.copy(order="F")


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def pzolopd1_device(dA, alpha: float, beta: float, rk: int, sym: str = "N", polar: bool = True,
                    precision: str = "f64", dU=None, dH=None):
    """PZOLOPD1 on device buffers (column-major, ld = n).  Returns (dU, dH, RunInfo)."""
    torch = _torch()
    n = dA.shape[0]
    if dU is None:
        dU = torch.empty_like(dA)
    if polar and dH is None:
        dH = torch.empty_like(dA)
    ci = _Info()
    st = _lib.zolo_pd_device(PREC[precision], sym.encode()[:1], rk, n, alpha, beta, C.c_void_p(dA.data_ptr()), n,
                             C.c_void_p(dU.data_ptr()), n, C.c_void_p(dH.data_ptr()) if polar else None, n,
                             C.byref(ci), _stream())
    _check("zolo_pd_device", st)
    return dU, (dH if polar else None), RunInfo.from_c(ci)


def pzolopd_auto_device(dA, precision: str = "f64", rk: int = 0, sym: str = "N", polar: bool = True, dU=None,
                        dH=None):
    """Zolo-PD with alpha, beta from the QR of A/alpha that the first QR step reuses
    (zolo_pd_auto_device).  Returns (dU, dH, RunInfo); RunInfo has alpha, beta, r."""
    torch = _torch()
    n = dA.shape[0]
    if dU is None:
        dU = torch.empty_like(dA)
    if polar and dH is None:
        dH = torch.empty_like(dA)
    if not polar:
        dH = None
    ci = _Info()
    st = _lib.zolo_pd_auto_device(PREC[precision], sym.encode()[:1], rk, n, C.c_void_p(dA.data_ptr()), n,
                                  C.c_void_p(dU.data_ptr()), n, C.c_void_p(dH.data_ptr()) if polar else None, n,
                                  C.byref(ci), _stream())
    _check("zolo_pd_auto_device", st)
    return dU, dH, RunInfo.from_c(ci)


def pqdwhfacs_device(dA, sigma: float = 0.0, polar: bool = True, sym: str = "N"):
    torch = _torch()
    n = dA.shape[0]
    dU = torch.empty_like(dA)
    dH = torch.empty_like(dA) if polar else None
    ci = _Info()
    st = _lib.zolo_qdwh_device(sym.encode()[:1], n, sigma, C.c_void_p(dA.data_ptr()), n, C.c_void_p(dU.data_ptr()), n,
                               C.c_void_p(dH.data_ptr()) if polar else None, n, C.byref(ci), _stream())
    _check("zolo_qdwh_device", st)
    return dU, dH, RunInfo.from_c(ci)


def bounds_device(dA, convention: str = "zolo"):
    """alpha = ||A||_F and the lower bound beta (test_zolodrv1.f90:179-189 / pqdwhfacs.f90:318-344)."""
    n = dA.shape[0]
    a, b = C.c_double(), C.c_double()
    conv = {"zolo": 0, "qdwh": 1}[convention]
    _check("zolo_bounds_device", _lib.zolo_bounds_device(conv, n, C.c_void_p(dA.data_ptr()), n, C.byref(a), C.byref(b),
                                                         _stream()))
    return a.value, b.value


def pdnormest_device(dA, tol: float = 1e-6):
    """PDNORMEST: power-iteration estimate of ||A||_2; returns (estimate, iterations)."""
    n, m = dA.shape  # column-major m x n stored as row-major n x m
    est, cnt = C.c_double(), C.c_int()
    _check("zolo_normest_device", _lib.zolo_normest_device(m, n, C.c_void_p(dA.data_ptr()), m, tol, C.byref(est),
                                                           C.byref(cnt), _stream()))
    return est.value, cnt.value


def gemm_device(transa: int, transb: int, m: int, n: int, k: int, alpha: float, dA, lda: int, dB, ldb: int,
                beta: float, dC, ldc: int, precision: str = "f64"):
    """C = alpha op(A) op(B) + beta C on device buffers (column-major).  precision: "f64", "f32",
    "bf16x3" or "f64x" (FP64 data, exact int8 products modulo 16 moduli)."""
    _check("zolo_gemm_device", _lib.zolo_gemm_device(GEMM_PREC[precision], transa, transb, m, n, k, alpha,
                                                     C.c_void_p(dA.data_ptr()), lda, C.c_void_p(dB.data_ptr()), ldb,
                                                     beta, C.c_void_p(dC.data_ptr()), ldc, _stream()))


def stacked_qr_device(dS, n: int, dQ=None, dtau=None):
    """MPDGEQRF + MPDORGQR of the 2n x n stack in dS (column-major, ld 2n)."""
    _check("zolo_stacked_qr_device", _lib.zolo_stacked_qr_device(
        n, C.c_void_p(dS.data_ptr()), 2 * n, C.c_void_p(dQ.data_ptr()) if dQ is not None else None, 2 * n,
        C.c_void_p(dtau.data_ptr()) if dtau is not None else None, _stream()))


def chol_inv_device(dA, dLi, precision: str = "f64"):
    n = dA.shape[0]
    _check("zolo_chol_inv_device", _lib.zolo_chol_inv_device(PREC[precision], n, C.c_void_p(dA.data_ptr()), n,
                                                             C.c_void_p(dLi.data_ptr()), n, _stream()))


class _BatchInfo(C.Structure):
    _fields_ = [
        ("count", C.c_int), ("buckets", C.c_int), ("fallbacks", C.c_int), ("failures", C.c_int),
        ("max_iters", C.c_int), ("graph_nodes", C.c_int), ("continued", C.c_int), ("time_ms", C.c_double),
        ("phase_ms", C.c_double * 8), ("flops_f32", C.c_double), ("flops_f64", C.c_double),
    ]


BATCH_PHASES = ("bounds", "plan", "tiles", "iterations", "fallback", "learned")
L0_SOURCE = {"lu": 0, "learned": 1, "given": 2}


@dataclass
class BatchRunInfo:
    """Per-matrix results and phase times of one batched call (zolo_batch_info_t)."""
    beta: np.ndarray
    orth: np.ndarray
    iters: np.ndarray
    fallback: np.ndarray
    failed: np.ndarray
    continued: np.ndarray
    buckets: int
    graph_nodes: int
    time_ms: float
    phase_ms: dict
    flops_f32: float = 0.0      # algorithmic flops of the executed schedules (zolo.h)
    flops_f64: float = 0.0


def pd_batched_device(mats, precision: str = "f32", l0: str = "lu", beta=None, orth_tol: float = 1e-6, rk: int = 0,
                      tile: int = -1, use_graph: bool = True, out=None):
    """Polar factors of many small FP32 matrices (orders 1..512) in one captured hipGraph.

    ``mats`` are float32 CUDA tensors holding column-major matrices (``to_device``
    layout: the row-major n x n storage of A^T).  ``l0`` selects the lower bound:
    ``"lu"`` (FP64 LU + 1-norm estimate, test_zolodrv1.f90:179-189), ``"learned"``
    (the network set by ``l0_set_weights``) or ``"given"`` (``beta``, one per
    matrix); the last two fall back to ``"lu"`` for a matrix that fails the
    orthogonality check.  Returns (list of U tensors, BatchRunInfo).
    """
    torch = _torch()
    k = len(mats)
    for m in mats:
        if m.dtype != torch.float32 or not m.is_cuda or m.dim() != 2 or m.shape[0] != m.shape[1]:
            raise ValueError("pd_batched_device: every matrix must be a square float32 CUDA tensor")
        if not m.is_contiguous():
            raise ValueError("pd_batched_device: matrices must be contiguous")
    outs = out if out is not None else [torch.empty_like(m) for m in mats]
    n = np.array([m.shape[0] for m in mats], dtype=np.int64)
    pa = np.array([m.data_ptr() for m in mats], dtype=np.uint64)
    pu = np.array([u.data_ptr() for u in outs], dtype=np.uint64)
    ld = n.copy()
    b_in = None
    if l0 == "given":
        b_in = np.ascontiguousarray(beta, dtype=np.float64)
        if b_in.shape != (k,):
            raise ValueError("pd_batched_device: beta needs one value per matrix")
    b_out = np.zeros(k)
    o_out = np.zeros(k)
    it_out = np.zeros(k, dtype=np.int32)
    fl_out = np.zeros(k, dtype=np.int32)
    ci = _BatchInfo()

    def ptr(a):
        return a.ctypes.data_as(C.c_void_p) if a is not None else None

    st = _lib.zolo_pd_batched_device(k, ptr(n), ptr(pa), ptr(ld), ptr(pu), ptr(ld), PREC[precision], L0_SOURCE[l0],
                                     ptr(b_in), orth_tol, rk, tile, 1 if use_graph else 0, ptr(b_out), ptr(o_out),
                                     ptr(it_out), ptr(fl_out), C.byref(ci), _stream())
    _check("zolo_pd_batched_device", st)
    info = BatchRunInfo(beta=b_out, orth=o_out, iters=it_out, fallback=(fl_out & 1) != 0, failed=(fl_out & 2) != 0,
                        continued=(fl_out & 4) != 0,
                        buckets=ci.buckets, graph_nodes=ci.graph_nodes, time_ms=ci.time_ms,
                        phase_ms={p: ci.phase_ms[i] for i, p in enumerate(BATCH_PHASES)},
                        flops_f32=ci.flops_f32, flops_f64=ci.flops_f64)
    return outs, info


def batched_tile(n: int, precision: str = "f32") -> int:
    """GEMM tile chosen by the autotuner for matrices of order n (-1: not tuned yet)."""
    return int(_lib.zolo_batched_tile(n, PREC[precision]))


# ---------------------------------------------------------------------------
# Zolo-SVD tail: symmetric eigensolver and the SVD assembly
# ---------------------------------------------------------------------------
def syevd_device(dH):
    """H = Z diag(w) Z^T on the device (column-major FP64, H overwritten); returns (w ascending, Z, phase_ms)."""
    torch = _torch()
    n = dH.shape[0]
    w = torch.empty((n,), dtype=torch.float64, device="cuda")
    Z = torch.empty_like(dH)
    ph = np.zeros(3)
    _check("zolo_syevd_device", _lib.zolo_syevd_device(n, C.c_void_p(dH.data_ptr()), n, C.c_void_p(w.data_ptr()),
                                                       C.c_void_p(Z.data_ptr()), n, ph.ctypes.data_as(C.c_void_p),
                                                       _stream()))
    return w, Z, dict(zip(("tridiagonal", "divide_conquer", "back_transform"), ph.tolist()))


BOUNDS = {"qr": 0, "lu": 1}


def svd_device(dA, precision: str = "f64", rk: int = 0, bounds: str = "qr"):
    """Zolo-SVD (ZoloCHK1.f90:80-91): A = U diag(S) V^T with S ascending; returns (S, U, V, RunInfo, phase_ms).
    bounds: "qr" (from the QR the first step reuses) or "lu" (LU + 1-norm estimate, test_zolodrv1.f90:179-189)."""
    torch = _torch()
    n = dA.shape[0]
    S = torch.empty((n,), dtype=torch.float64, device="cuda")
    U = torch.empty_like(dA)
    V = torch.empty_like(dA)
    ci = _Info()
    ph = np.zeros(5)
    _check("zolo_svd_ex_device", _lib.zolo_svd_ex_device(PREC[precision], BOUNDS[bounds], rk, n,
                                                         C.c_void_p(dA.data_ptr()), n, C.c_void_p(S.data_ptr()),
                                                         C.c_void_p(U.data_ptr()), n, C.c_void_p(V.data_ptr()), n,
                                                         C.byref(ci), ph.ctypes.data_as(C.c_void_p), _stream()))
    names = ("polar_h", "tridiagonal", "divide_conquer", "back_transform", "eig_total")
    return S, U, V, RunInfo.from_c(ci), dict(zip(names, ph.tolist()))


MM_TYPES = {0: "symmetric", 1: "general", 2: "other"}


def mm_info(path: str, bandwidth: bool = False) -> dict:
    """Header of a MatrixMarket coordinate file (and max |i - j| + 1 when asked: reads every entry)."""
    r, c, z, bw = _I64(0), _I64(0), _I64(0), _I64(0)
    t = C.c_int(0)
    _check("zolo_mm_info", _lib.zolo_mm_info(os.fsencode(path), C.byref(r), C.byref(c), C.byref(z), C.byref(t),
                                             C.byref(bw) if bandwidth else None))
    out = dict(rows=r.value, cols=c.value, nnz=z.value, mtype=t.value, kind=MM_TYPES[t.value])
    if bandwidth:
        out["bandwidth"] = bw.value
    return out


def formx2(X, Z, fone: float):
    """FormX2 (FormX2.f90:2) on the GPU: returns (X+ = (X + Z)/FONE, ERR = ||X+ - X||_F)."""
    A = _fortran(X).copy(order="F")
    Zc = _fortran(Z).copy(order="F")
    n = A.shape[0]
    err = C.c_double(0.0)
    _check("zolo_formx2_host", _lib.zolo_formx2_host(n, float(fone), _dptr(A), n, _dptr(Zc), n, C.byref(err)))
    return A, err.value


def mm_read(path: str, sigma: float = 0.0):
    """Dense matrix of a MatrixMarket file on the host, assembled as prepare_real_GEM: symmetric
    files get each stored entry on both sides of the diagonal and the diagonal shifted by -sigma
    (A <- A + A^T, then A(i, i)/2 - sigma); other types are kept as stored.
    Returns (A Fortran-ordered, mtype)."""
    n = mm_info(path)["rows"]
    A = np.zeros((n, n), order="F")
    t = C.c_int(0)
    _check("zolo_mm_dense_host", _lib.zolo_mm_dense_host(os.fsencode(path), float(sigma), _dptr(A), max(n, 1),
                                                         C.byref(t)))
    return A, t.value


def mm_read_device(path: str, sigma: float = 0.0):
    """As ``mm_read``, assembled on the GPU from the uploaded entries; returns (column-major tensor, mtype)."""
    torch = _torch()
    n = mm_info(path)["rows"]
    dA = torch.empty((n, n), dtype=torch.float64, device="cuda")
    t = C.c_int(0)
    _check("zolo_mm_dense_device", _lib.zolo_mm_dense_device(os.fsencode(path), float(sigma),
                                                             C.c_void_p(dA.data_ptr()), max(n, 1), C.byref(t),
                                                             _stream()))
    return dA, t.value


def subit_device(dA, normA: float, dQ):
    """PQDWHSUBIT (pqdwhsubit.f90:2): dQ the polar factor of A - sigma I on entry,
    V = [V1 V2] on exit; dA overwritten by V^T A V.  Returns (rk, err, iterations)."""
    n = dA.shape[0]
    rk, it = C.c_int(0), C.c_int(0)
    err = C.c_double(0.0)
    _check("zolo_subit_device", _lib.zolo_subit_device(n, C.c_void_p(dA.data_ptr()), n, float(normA),
                                                       C.c_void_p(dQ.data_ptr()), n, C.byref(rk), C.byref(err),
                                                       C.byref(it), _stream()))
    return rk.value, err.value, it.value


def eig_device(dQ, dAs):
    """ZoloEig (ZoloEig.f90:2): dQ the polar factor of As - sigma I, dAs the matrix
    (both overwritten).  Returns (w, V, rk, err, iterations): the rk eigenvalues
    below sigma (ascending) and V column-major n x rk (a torch (rk, n) tensor)."""
    torch = _torch()
    n = dQ.shape[0]
    w = torch.zeros((n,), dtype=torch.float64, device="cuda")
    V = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    rk, it = C.c_int(0), C.c_int(0)
    err = C.c_double(0.0)
    _check("zolo_eig_device", _lib.zolo_eig_device(n, C.c_void_p(dQ.data_ptr()), n, C.c_void_p(dAs.data_ptr()), n,
                                                   C.c_void_p(w.data_ptr()), C.c_void_p(V.data_ptr()), n,
                                                   C.byref(rk), C.byref(err), C.byref(it), _stream()))
    return w[:rk.value], V[:rk.value], rk.value, err.value, it.value


def eig_below_device(dA, sigma: float, precision: str = "f64", rk: int = 0, bounds: str = "qr"):
    """Eigenpairs of the symmetric A below sigma (pqdwheig.f90:113-144 with the Zolo
    polar factor of A - sigma I).  dA is not modified.  Returns
    (w, V, count, RunInfo, err, iterations) as ``eig_device``.  bounds: "qr" (from the
    QR the first step reuses) or "lu" (LU + 1-norm estimate of A - sigma I with
    PQDWHFACS's convention, pqdwhfacs.f90:318-344)."""
    torch = _torch()
    n = dA.shape[0]
    w = torch.zeros((n,), dtype=torch.float64, device="cuda")
    V = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    cnt, it = C.c_int(0), C.c_int(0)
    err = C.c_double(0.0)
    ci = _Info()
    _check("zolo_eig_below_ex_device",
           _lib.zolo_eig_below_ex_device(PREC[precision], BOUNDS[bounds], rk, n, C.c_void_p(dA.data_ptr()), n,
                                         float(sigma), C.c_void_p(w.data_ptr()), C.c_void_p(V.data_ptr()), n,
                                         C.byref(cnt), C.byref(ci), C.byref(err), C.byref(it), _stream()))
    k = cnt.value
    return w[:k], V[:k], k, RunInfo.from_c(ci), err.value, it.value


# ---------------------------------------------------------------------------
# Learned lower bound (features, network, training on the device)
# ---------------------------------------------------------------------------
L0_WEIGHTS = os.path.join(_HERE, "l0_weights.json")


def l0_param_count() -> int:
    return int(_lib.zolo_l0_param_count())


def l0_set_weights(params) -> None:
    """Install the network parameters (zolo_l0_param_count() floats) for ``l0="learned"``."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    _check("zolo_l0_set_weights", _lib.zolo_l0_set_weights(p.ctypes.data_as(C.c_void_p), int(p.size)))


def l0_load_weights(path: str = L0_WEIGHTS) -> dict:
    """Load a weights file written by ``zolo.train_l0`` and install it; returns its metadata."""
    import json
    with open(path) as f:
        meta = json.load(f)
    l0_set_weights(np.array(meta["params"], dtype=np.float32))
    return meta


def _mat_arrays(mats):
    n = np.array([m.shape[0] for m in mats], dtype=np.int64)
    pa = np.array([m.data_ptr() for m in mats], dtype=np.uint64)
    return n, pa, n.copy()


def l0_features_device(mats):
    """Features (count x 16 float32 CUDA tensor) and smax (count) of column-major float32 matrices."""
    torch = _torch()
    k = len(mats)
    feat = torch.empty((k, 16), dtype=torch.float32, device="cuda")
    smax = torch.empty((k,), dtype=torch.float32, device="cuda")
    n, pa, ld = _mat_arrays(mats)
    _check("zolo_l0_features_device", _lib.zolo_l0_features_device(
        k, n.ctypes.data_as(C.c_void_p), pa.ctypes.data_as(C.c_void_p), ld.ctypes.data_as(C.c_void_p),
        C.c_void_p(feat.data_ptr()), C.c_void_p(smax.data_ptr()), _stream()))
    return feat, smax


def l0_forward_device(params, feat):
    """Network output y = predicted log10(sigma_min/sigma_max) for each feature row."""
    torch = _torch()
    y = torch.empty((feat.shape[0],), dtype=torch.float32, device="cuda")
    _check("zolo_l0_forward_device", _lib.zolo_l0_forward_device(
        C.c_void_p(params.data_ptr()), C.c_void_p(feat.data_ptr()), feat.shape[0], C.c_void_p(y.data_ptr()),
        _stream()))
    return y


def l0_grad_device(params, feat, target, tau: float):
    """(gradient of the pinball loss w.r.t. the trained block, loss) computed by the device kernels."""
    torch = _torch()
    g = torch.empty((l0_param_count() - 33,), dtype=torch.float32, device="cuda")
    loss = C.c_float()
    _check("zolo_l0_grad_device", _lib.zolo_l0_grad_device(
        C.c_void_p(params.data_ptr()), C.c_void_p(feat.data_ptr()), C.c_void_p(target.data_ptr()), feat.shape[0],
        tau, C.c_void_p(g.data_ptr()), C.byref(loss), _stream()))
    return g, loss.value


def l0_train_device(params, feat, target, steps: int, lr: float = 1e-3, tau: float = 0.05):
    """Full-batch Adam on the device, params (float32 CUDA tensor) updated in place; returns the loss history."""
    hist = np.zeros(max(steps, 1), dtype=np.float32)
    _check("zolo_l0_train_device", _lib.zolo_l0_train_device(
        C.c_void_p(params.data_ptr()), C.c_void_p(feat.data_ptr()), C.c_void_p(target.data_ptr()), feat.shape[0],
        steps, lr, tau, hist.ctypes.data_as(C.c_void_p), _stream()))
    return hist[:steps]


def finalize():
    _check("zolo_finalize", _lib.zolo_finalize())
