// extern "C" entry points of libzolo.so (declared in include/zolo.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "batched.hpp"
#include "mlp.hpp"
#include "coeff.hpp"
#include "driver.hpp"
#include "eig.hpp"

using namespace zolo;

namespace {

// Copies host column-major matrices to the device and back for the host-pointer entry points.
struct DevMat {
  double* p = nullptr;
  ~DevMat() {
    if (p) (void)hipFree(p);
  }
  int alloc(int64_t n) {
    if (hipMalloc(&p, sizeof(double) * (size_t)n * n) != hipSuccess) {
      p = nullptr;
      set_error("host entry: hipMalloc failed");
      return kErrAlloc;
    }
    return kOk;
  }
  int upload(const double* h, int64_t ld, int64_t n) {
    ZOLO_HIP(hipMemcpy2D(p, n * sizeof(double), h, ld * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice));
    return kOk;
  }
  int download(double* h, int64_t ld, int64_t n) {
    ZOLO_HIP(hipMemcpy2D(h, ld * sizeof(double), p, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost));
    return kOk;
  }
};

int g_device = -1;

int ensure_device() {
  if (g_device < 0) {
    ZOLO_HIP(hipGetDevice(&g_device));
  }
  return kOk;
}

// Every extern "C" entry point below is a function-try-block: an exception
// from the host side (a failed host allocation, a library error) becomes a
// status code instead of leaving the C ABI.
int exception_status() {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return kErrAlloc;
  } catch (const std::exception& e) {
    set_error(std::string("internal error: ") + e.what());
    return kErrInternal;
  } catch (...) {
    set_error("internal error");
    return kErrInternal;
  }
}

}  // namespace

extern "C" {

int zolo_init(int device) try {
  if (device >= 0) ZOLO_HIP(hipSetDevice(device));
  ZOLO_HIP(hipGetDevice(&g_device));
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_finalize(void) try {
  workspace_release_all();
  l0_release();
  return kOk;
} catch (...) {
  return exception_status();
}

const char* zolo_last_error(void) { return last_error(); }

int zolo_version(void) { return 100; }

int zolo_compute_coeff(double con, double tol, int rk, double* coeff) try {
  if (!(con >= 1.0)) return -1;
  if (rk < 1 || rk > 8) return -3;
  compute_coeff(con, tol, rk, coeff);
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_choosem(double con) { return choosem(con); }
double zolo_hfunc(double x) { return hfunc(x); }
void zolo_mellipke(double alpha, double tol, double* k, double* e) { mellipke(alpha, tol, k, e); }
void zolo_mellipj(double u, double alpha, double tol, double* sn, double* cn, double* dn) {
  mellipj(u, alpha, tol, sn, cn, dn);
}

// Argument checks of the device-pointer Zolo-PD entry points, in the argument
// positions of zolo_pd_device; they run before any HIP call.
static int pd_device_args(int precision, char sym, int rk, int64_t n, const double* A, int64_t lda, const double* U,
                          int64_t ldu, const double* H, int64_t ldh, bool autob) {
  if (precision < 0 || precision > 4) return -1;
  if (sym != 'S' && sym != 's' && sym != 'N' && sym != 'n') return -2;
  if (rk < (autob ? 0 : 1) || rk > 8) return -3;
  if (n < 1) return -4;
  if (!A) return -7;
  if (lda < n) return -8;
  if (!U) return -9;
  if (ldu < n) return -10;
  if (H && ldh < n) return -12;
  return kOk;
}

int zolo_pd_device(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A,
                   int64_t lda, double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info, void* stream) try {
  if (!(alpha > 0.0)) return -5;  // (alpha, beta) = (0, 0) is zolo_pd_auto_device's mode, not this one's
  if (!(beta > 0.0) || !(beta <= 1.0)) return -6;
  ZOLO_CHECK(pd_device_args(precision, sym, rk, n, A, lda, U, ldu, H, ldh, false));
  ZOLO_CHECK(ensure_device());
  return zolo_pd_run(precision, sym, rk, n, alpha, beta, A, lda, U, ldu, H, ldh, info, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_pd_auto_device(int precision, char sym, int rk, int64_t n, const double* A, int64_t lda, double* U,
                        int64_t ldu, double* H, int64_t ldh, zolo_info_t* info, void* stream) try {
  // zolo_pd_device's positions 7..12 are this function's 5..10 (no alpha, beta)
  const int st = pd_device_args(precision, sym, rk, n, A, lda, U, ldu, H, ldh, true);
  if (st < 0) return st <= -7 ? st + 2 : st;
  ZOLO_CHECK(ensure_device());
  return zolo_pd_run(precision, sym, rk, n, 0.0, 0.0, A, lda, U, ldu, H, ldh, info, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_pd_host(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A,
                 int64_t lda, double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info) try {
  if (n < 1) return -4;
  if (!(alpha > 0.0)) return -5;
  if (!(beta > 0.0) || !(beta <= 1.0)) return -6;
  if (lda < n) return -8;
  if (ldu < n) return -10;
  if (H && ldh < n) return -12;
  ZOLO_CHECK(ensure_device());
  DevMat dA, dU, dH;
  ZOLO_CHECK(dA.alloc(n));
  ZOLO_CHECK(dU.alloc(n));
  if (H) ZOLO_CHECK(dH.alloc(n));
  ZOLO_CHECK(dA.upload(A, lda, n));
  const int st = zolo_pd_run(precision, sym, rk, n, alpha, beta, dA.p, n, dU.p, n, H ? dH.p : nullptr, n, info, nullptr);
  if (st < 0) return st;
  ZOLO_HIP(hipDeviceSynchronize());
  ZOLO_CHECK(dU.download(U, ldu, n));
  if (H) ZOLO_CHECK(dH.download(H, ldh, n));
  return st;
} catch (...) {
  return exception_status();
}

int zolo_pd2_host(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A,
                  int64_t lda, double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info) try {
  if (n < 1) return -4;
  if (!(alpha > 0.0)) return -5;
  if (!(beta > 0.0) || !(beta <= 1.0)) return -6;
  if (lda < n) return -8;
  if (ldu < n) return -10;
  if (H && ldh < n) return -12;
  ZOLO_CHECK(ensure_device());
  DevMat dA, dU, dH;
  ZOLO_CHECK(dA.alloc(n));
  ZOLO_CHECK(dU.alloc(n));
  if (H) ZOLO_CHECK(dH.alloc(n));
  ZOLO_CHECK(dA.upload(A, lda, n));
  const int st = zolo_pd_run(precision, sym, rk, n, alpha, beta, dA.p, n, dU.p, n, H ? dH.p : nullptr, n, info,
                             nullptr, 12, true);
  if (st < 0) return st;
  ZOLO_HIP(hipDeviceSynchronize());
  ZOLO_CHECK(dU.download(U, ldu, n));
  if (H) ZOLO_CHECK(dH.download(H, ldh, n));
  return st;
} catch (...) {
  return exception_status();
}

int zolo_bounds_host(int convention, int64_t n, const double* A, int64_t lda, double* alpha, double* beta) try {
  if (convention != 0 && convention != 1) return -1;
  if (n < 1) return -2;
  if (lda < n) return -4;
  ZOLO_CHECK(ensure_device());
  DevMat dA;
  ZOLO_CHECK(dA.alloc(n));
  ZOLO_CHECK(dA.upload(A, lda, n));
  const size_t bytes = bounds_bytes(n);
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  const int st = bounds_run(convention, n, dA.p, n, alpha, beta, ar, nullptr);
  ZOLO_HIP(hipDeviceSynchronize());
  return st;
} catch (...) {
  return exception_status();
}

int zolo_normest_host(int64_t m, int64_t n, const double* A, int64_t lda, double tol, double* est, int* cnt) try {
  if (m < 1) return -1;
  if (n < 1) return -2;
  if (lda < m) return -4;
  ZOLO_CHECK(ensure_device());
  double* dA = nullptr;
  if (hipMalloc(&dA, sizeof(double) * (size_t)m * n) != hipSuccess) {
    set_error("normest: hipMalloc failed");
    return kErrAlloc;
  }
  int st = kOk;
  if (hipMemcpy2D(dA, m * sizeof(double), A, lda * sizeof(double), m * sizeof(double), n, hipMemcpyHostToDevice) !=
      hipSuccess) {
    set_error("normest: copy failed");
    st = kErrHip;
  }
  if (st == kOk) {
    const size_t bytes = normest_bytes(m, n);
    char* base = nullptr;
    st = workspace_acquire(bytes, &base);
    if (st == kOk) {
      Arena ar(base, bytes);
      st = normest_run(m, n, dA, m, tol, est, cnt, ar, nullptr);
    }
  }
  (void)hipDeviceSynchronize();
  (void)hipFree(dA);
  return st;
} catch (...) {
  return exception_status();
}

// Host stacked QR pieces: the 2n x n stack S (ld lds) is copied to the device.
static int stacked_host(int64_t n, double* S, int64_t lds, double* tau, bool geqrf) {
  if (n < 1) return -1;
  if (lds < 2 * n) return -3;
  ZOLO_CHECK(ensure_device());
  const int64_t ls = 2 * n;
  double *dS = nullptr, *dQ = nullptr, *dt = nullptr;
  int st = kOk;
  if (hipMalloc(&dS, sizeof(double) * (size_t)ls * n) != hipSuccess ||
      hipMalloc(&dQ, sizeof(double) * (size_t)ls * n) != hipSuccess ||
      hipMalloc(&dt, sizeof(double) * (size_t)n) != hipSuccess) {
    set_error("stacked qr: hipMalloc failed");
    st = kErrAlloc;
  }
  if (st == kOk && (hipMemcpy2D(dS, ls * sizeof(double), S, lds * sizeof(double), ls * sizeof(double), n,
                                hipMemcpyHostToDevice) != hipSuccess ||
                    (!geqrf && hipMemcpy(dt, tau, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess))) {
    set_error("stacked qr: copy failed");
    st = kErrHip;
  }
  if (st == kOk) {
    const size_t bytes = qr_step_bytes<double>(n, 1);
    char* base = nullptr;
    st = workspace_acquire(bytes, &base);
    if (st == kOk) {
      Arena ar(base, bytes);
      Ctx c;
      st = geqrf ? stacked_geqrf<double>(c, dS, ls, 0, n, 1, dt, ar)
                 : stacked_orgqr<double>(c, dS, ls, 0, dt, dQ, ls, 0, n, 1, ar);
    }
  }
  if (st == kOk) {
    const double* src = geqrf ? dS : dQ;
    if (hipMemcpy2D(S, lds * sizeof(double), src, ls * sizeof(double), ls * sizeof(double), n,
                    hipMemcpyDeviceToHost) != hipSuccess ||
        (geqrf && hipMemcpy(tau, dt, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)) {
      set_error("stacked qr: copy back failed");
      st = kErrHip;
    }
  }
  (void)hipFree(dS);
  (void)hipFree(dQ);
  (void)hipFree(dt);
  return st;
}

int zolo_mpdgeqrf_host(int64_t n, double* S, int64_t lds, double* tau) { return stacked_host(n, S, lds, tau, true); }
int zolo_mpdorgqr_host(int64_t n, double* S, int64_t lds, const double* tau) try {
  return stacked_host(n, S, lds, const_cast<double*>(tau), false);
} catch (...) {
  return exception_status();
}

int zolo_bounds_device(int convention, int64_t n, const double* A, int64_t lda, double* alpha, double* beta,
                       void* stream) try {
  if (convention != 0 && convention != 1) return -1;
  if (n < 1) return -2;
  if (lda < n) return -4;
  ZOLO_CHECK(ensure_device());
  const size_t bytes = bounds_bytes(n);
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  return bounds_run(convention, n, A, lda, alpha, beta, ar, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_qdwh_device(char sym, int64_t n, double sigma, const double* A, int64_t lda, double* U, int64_t ldu,
                     double* H, int64_t ldh, zolo_info_t* info, void* stream) try {
  if (sym != 'S' && sym != 's' && sym != 'N' && sym != 'n') return -1;
  if (n < 1) return -2;
  if (!std::isfinite(sigma)) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!U) return -6;
  if (ldu < n) return -7;
  if (H && ldh < n) return -9;
  ZOLO_CHECK(ensure_device());
  return qdwh_run(sym, n, sigma, A, lda, U, ldu, H, ldh, info, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_qdwh_host(char sym, int64_t n, double sigma, const double* A, int64_t lda, double* U, int64_t ldu, double* H,
                   int64_t ldh, zolo_info_t* info) try {
  if (n < 1) return -2;
  if (lda < n) return -5;
  if (ldu < n) return -7;
  if (H && ldh < n) return -9;
  ZOLO_CHECK(ensure_device());
  DevMat dA, dU, dH;
  ZOLO_CHECK(dA.alloc(n));
  ZOLO_CHECK(dU.alloc(n));
  if (H) ZOLO_CHECK(dH.alloc(n));
  ZOLO_CHECK(dA.upload(A, lda, n));
  const int st = qdwh_run(sym, n, sigma, dA.p, n, dU.p, n, H ? dH.p : nullptr, n, info, nullptr);
  if (st < 0) return st;
  ZOLO_HIP(hipDeviceSynchronize());
  ZOLO_CHECK(dU.download(U, ldu, n));
  if (H) ZOLO_CHECK(dH.download(H, ldh, n));
  return st;
} catch (...) {
  return exception_status();
}

int zolo_chol_step_host(char sym, int64_t n, const double* X, int64_t ldx, double* Z, int64_t ldz, const double* coeff,
                        int rk, int ip) try {
  if (sym != 'S' && sym != 's' && sym != 'N' && sym != 'n') return -1;
  if (n < 1) return -2;
  if (ldx < n) return -4;
  if (ldz < n) return -6;
  if (rk < 1 || rk > 8) return -8;
  if (ip < 1 || ip > rk) return -9;
  ZOLO_CHECK(ensure_device());
  DevMat dX, dZ;
  ZOLO_CHECK(dX.alloc(n));
  ZOLO_CHECK(dZ.alloc(n));
  ZOLO_CHECK(dX.upload(X, ldx, n));
  const size_t bytes = chol_step_bytes<double>(n, 1) + 8 * 256;
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  double* dsc = ar.take<double>(4);
  int* dinfo = ar.take<int>(4);
  const double hs[2] = {coeff[2 * ip - 2], -coeff[2 * rk + ip - 1]};
  ZOLO_HIP(hipMemcpy(dsc, hs, sizeof(hs), hipMemcpyHostToDevice));
  ZOLO_HIP(hipMemset(dinfo, 0, sizeof(int)));
  Ctx c;
  const bool left = (sym == 'S' || sym == 's');
  ZOLO_CHECK(chol_step<double>(c, dX.p, n, n, 1, dsc, 1.0, 0.0, dsc + 1, dZ.p, n, n * n, ar, dinfo, left));
  int hinfo = 0;
  ZOLO_HIP(hipMemcpy(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost));
  ZOLO_CHECK(dZ.download(Z, ldz, n));
  if (hinfo) {
    set_error("zolo_chol_step: not positive definite");
    return kErrNotSpd;
  }
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_qr_step_host(int64_t n, const double* X, int64_t ldx, double* Z, int64_t ldz, const double* coeff, int rk,
                      int ip) try {
  if (n < 1) return -1;
  if (ldx < n) return -3;
  if (ldz < n) return -5;
  if (rk < 1 || rk > 8) return -7;
  if (ip < 1 || ip > rk) return -8;
  ZOLO_CHECK(ensure_device());
  DevMat dX, dZ;
  ZOLO_CHECK(dX.alloc(n));
  ZOLO_CHECK(dZ.alloc(n));
  ZOLO_CHECK(dX.upload(X, ldx, n));
  const size_t bytes = qr_step_bytes<double>(n, 1) + 8 * 256;
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  double* dsc = ar.take<double>(4);
  const double c = coeff[2 * ip - 2];
  const double hs[2] = {std::sqrt(c), -coeff[2 * rk + ip - 1] / std::sqrt(c)};
  ZOLO_HIP(hipMemcpy(dsc, hs, sizeof(hs), hipMemcpyHostToDevice));
  Ctx cx;
  int nz = 1;
  ZOLO_CHECK(qr_step<double>(cx, dX.p, n, n, 1, dsc, dsc + 1, dZ.p, n, n * n, ar, &nz));
  ZOLO_HIP(hipDeviceSynchronize());
  ZOLO_CHECK(dZ.download(Z, ldz, n));
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_normest_device(int64_t m, int64_t n, const double* A, int64_t lda, double tol, double* est, int* cnt,
                        void* stream) try {
  if (m < 1) return -1;
  if (n < 1) return -2;
  if (lda < m) return -4;
  ZOLO_CHECK(ensure_device());
  const size_t bytes = normest_bytes(m, n);
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  return normest_run(m, n, A, lda, tol, est, cnt, ar, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_gemm_device(int precision, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                     const void* A, int64_t lda, const void* B, int64_t ldb, double beta, void* C, int64_t ldc,
                     void* stream) try {
  if (precision < 0 || precision > 3) return -1;
  if (m < 0) return -4;
  if (n < 0) return -5;
  if (k < 0) return -6;
  if (lda < std::max<int64_t>(1, transa ? k : m)) return -9;
  if (ldb < std::max<int64_t>(1, transb ? n : k)) return -11;
  if (ldc < std::max<int64_t>(1, m)) return -14;
  GemmDesc d;
  d.transA = transa; d.transB = transb; d.M = m; d.N = n; d.K = k;
  d.A = A; d.lda = lda; d.B = B; d.ldb = ldb; d.C = C; d.ldc = ldc; d.alpha = alpha; d.beta = beta;
  return gemm((Prec)precision, d, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_set_f64_products(int mode) try {
  if (mode < 0 || mode > 1) return -1;
  set_f64_products(mode);
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_get_f64_products(void) { return f64_products(); }

int zolo_set_panel_exchange(int mode) try {
  if (mode < 0 || mode > 1) return -1;
  qr_set_panel_device_stores(mode);
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_get_panel_exchange(void) { return qr_panel_device_stores(); }

int zolo_profile_enable(int on) try {
  ledger_enable(on != 0);
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_profile_gemm(int precision, int transa, int transb, int tile_m, int tile_n, int64_t* launches, double* ms,
                      double* flops) try {
  return ledger_query(precision, transa, transb, tile_m, tile_n, launches, ms, flops);
} catch (...) {
  return exception_status();
}

int zolo_stacked_qr_device(int64_t n, double* S, int64_t lds, double* Q, int64_t ldq, double* tau, void* stream) try {
  if (n < 1) return -1;
  if (lds < 2 * n) return -3;
  if (Q && ldq < 2 * n) return -5;
  ZOLO_CHECK(ensure_device());
  const size_t bytes = qr_step_bytes<double>(n, 1);
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  Ctx c;
  c.s = (hipStream_t)stream;
  return stacked_qr<double>(c, S, lds, 0, Q, ldq, 0, n, 1, tau, ar);
} catch (...) {
  return exception_status();
}

int zolo_chol_inv_device(int precision, int64_t n, void* A, int64_t lda, void* Li, int64_t ldl, void* stream) try {
  if (precision < 0 || precision > 2) return -1;
  if (n < 1) return -2;
  if (lda < n) return -4;
  if (ldl < n) return -6;
  ZOLO_CHECK(ensure_device());
  const int64_t h = (n + 1) / 2 + kCholLeaf;
  const size_t bytes = 8 * (size_t)h * h + 8 * 256;
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  int* dinfo = ar.take<int>(4);
  Ctx c;
  c.s = (hipStream_t)stream;
  c.low = precision == 2 ? Prec::BF16X3 : Prec::F32;
  ZOLO_HIP(hipMemsetAsync(dinfo, 0, sizeof(int), c.s));
  int st;
  if (precision == 0) {
    double* tmp = ar.take<double>((size_t)h * h);
    st = chol_inv<double>(c, (double*)A, lda, 0, (double*)Li, ldl, 0, n, tmp, h * h, dinfo, 1);
  } else {
    float* tmp = ar.take<float>((size_t)h * h);
    st = chol_inv<float>(c, (float*)A, lda, 0, (float*)Li, ldl, 0, n, tmp, h * h, dinfo, 1);
  }
  if (st) return st;
  int hinfo = 0;
  ZOLO_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, c.s));
  ZOLO_HIP(hipStreamSynchronize(c.s));
  if (hinfo) {
    set_error("zolo_chol_inv: not positive definite at column " + std::to_string(hinfo));
    return kErrNotSpd;
  }
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_pd_batched_device(int count, const int64_t* n, const float* const* A, const int64_t* lda, float* const* U,
                           const int64_t* ldu, int precision, int l0_source, const double* beta_in, double orth_tol,
                           int rk, int tile, int use_graph, double* beta_out, double* orth_out, int* iters_out,
                           int* flags_out, zolo_batch_info_t* info, void* stream) try {
  if (count < 0) return -1;
  if (count > 0 && (!n || !A || !lda || !U || !ldu)) return -2;
  if (precision != 1) {
    set_error("batched PD: only FP32 iterations (precision 1) reach the 1e-6 orthogonality target");
    return -7;
  }
  if (l0_source < 0 || l0_source > 2) return -8;
  if (l0_source == 2 && !beta_in) return -9;
  if (!(orth_tol > 0.0)) return -10;
  if (rk < 0 || rk > 8) return -11;
  if (tile < -1 || tile > 4) return -12;
  ZOLO_CHECK(ensure_device());
  BatchOptions opt;
  opt.precision = precision;
  opt.orth_tol = orth_tol;
  opt.rk = rk;
  opt.tile = tile;
  opt.use_graph = use_graph != 0;
  std::vector<double> learned;
  double learned_ms = 0.0;
  if (l0_source == 1) {
    ZOLO_CHECK(l0_learned_beta(count, n, A, lda, learned, &learned_ms, (hipStream_t)stream));
    opt.l0_source = 2;
    opt.beta_in = learned.data();
  } else {
    opt.l0_source = l0_source;
    opt.beta_in = beta_in;
  }
  BatchResult res;
  const int st = pd_batched_run(count, n, A, lda, U, ldu, opt, res, (hipStream_t)stream);
  if (st < 0) {
    // argument numbers of pd_batched_run follow this function's own list
    return st;
  }
  if (st) return st;
  int fb = 0, fail = 0, mx = 0, cont = 0;
  for (int i = 0; i < count; ++i) {
    cont += res.continued[i] ? 1 : 0;
    if (beta_out) beta_out[i] = res.beta[i];
    if (orth_out) orth_out[i] = res.orth[i];
    if (iters_out) iters_out[i] = res.iters[i];
    if (flags_out) flags_out[i] = (res.fallback[i] ? 1 : 0) | (res.failed[i] ? 2 : 0) | (res.continued[i] ? 4 : 0);
    fb += res.fallback[i] ? 1 : 0;
    fail += res.failed[i] ? 1 : 0;
    mx = std::max(mx, res.iters[i]);
  }
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->count = count;
    info->buckets = res.buckets;
    info->fallbacks = fb;
    info->continued = cont;
    info->failures = fail;
    info->max_iters = mx;
    info->graph_nodes = (int)res.graph_nodes;
    for (int k = 0; k < 5; ++k) info->phase_ms[k] = res.phase_ms[k];
    info->phase_ms[5] = learned_ms;
    info->time_ms = res.time_ms + learned_ms;
    info->flops_f32 = res.flops_f32;
    info->flops_f64 = res.flops_f64;
  }
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_batched_tile(int64_t n, int precision) { return batched_tile_of(n, precision); }

int zolo_syevd_device(int64_t n, double* H, int64_t ldh, double* w, double* Z, int64_t ldz, double* phase_ms,
                      void* stream) try {
  if (n < 1 || n > 32768) return -1;
  if (!H) return -2;
  if (ldh < n) return -3;
  if (!w) return -4;
  if (!Z) return -5;
  if (ldz < n) return -6;
  ZOLO_CHECK(ensure_device());
  return syevd_run(n, H, ldh, w, Z, ldz, (hipStream_t)stream, phase_ms);
} catch (...) {
  return exception_status();
}

int zolo_subit_device(int64_t n, double* A, int64_t lda, double normA, double* Q, int64_t ldq, int* rk, double* err,
                      int* iters, void* stream) try {
  if (n < 1 || n > 32768) return -1;
  if (!A) return -2;
  if (lda < n) return -3;
  if (!(normA > 0.0)) return -4;
  if (!Q) return -5;
  if (ldq < n) return -6;
  if (!rk) return -7;
  ZOLO_CHECK(ensure_device());
  return subit_run(n, A, lda, normA, Q, ldq, rk, err, iters, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_eig_device(int64_t n, double* Q, int64_t ldq, double* As, int64_t lda, double* w, double* V, int64_t ldv,
                    int* rk, double* err, int* iters, void* stream) try {
  if (n < 1 || n > 32768) return -1;
  if (!Q) return -2;
  if (ldq < n) return -3;
  if (!As) return -4;
  if (lda < n) return -5;
  if (!w) return -6;
  if (!V) return -7;
  if (ldv < n) return -8;
  if (!rk) return -9;
  ZOLO_CHECK(ensure_device());
  return zoloeig_run(n, Q, ldq, As, lda, w, V, ldv, rk, err, iters, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

namespace {
// pqdwheig.f90:113-144 with the Zolo polar factor, on validated arguments
int eig_below_run(int precision, int bounds, int rk, int64_t n, const double* A, int64_t lda, double sigma, double* w,
                  double* V, int64_t ldv, int* count, zolo_info_t* info, double* err, int* iters, void* stream) {
  ZOLO_CHECK(ensure_device());
  hipStream_t s = (hipStream_t)stream;
  DevMat B, Q, As;
  ZOLO_CHECK(B.alloc(n));
  ZOLO_CHECK(Q.alloc(n));
  ZOLO_CHECK(As.alloc(n));
  // B = A - sigma I (pqdwhfacs with SIGMA, pqdwheig.f90:113-114); As keeps A
  ZOLO_HIP(hipMemcpy2DAsync(As.p, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                            hipMemcpyDeviceToDevice, s));
  ZOLO_HIP(hipMemcpy2DAsync(B.p, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                            hipMemcpyDeviceToDevice, s));
  if (sigma != 0.0) ZOLO_CHECK(k_add_diag<double>(B.p, n, n, -sigma, s));
  zolo_info_t local{};
  // polar factor of the symmetric A - sigma I into Q (rk = 0: choosem), its
  // bounds from the QR the first step reuses (bounds = 0), or by LU + 1-norm
  // estimate as PQDWHFACS computes them in the reference's pqdwheig flow
  // (bounds = 1, pqdwhfacs.f90:318-344: beta = ||B/alpha||_1 rcond / sqrt(n))
  double alpha = 0.0, beta = 0.0;
  if (bounds == 1) {
    ZOLO_CHECK(zolo_bounds_device(1, n, B.p, n, &alpha, &beta, stream));
    if (rk == 0) rk = choosem(1.0 / beta);
  }
  const int st = zolo_pd_run(precision, 'S', rk, n, alpha, beta, B.p, n, Q.p, n, nullptr, n, info ? info : &local, s);
  if (st) return st;
  // PQDWHSUBIT, PDSYEVD on the Rk block, V1 VV1 (pqdwheig.f90:119-144 = ZoloEig.f90:80-103)
  return zoloeig_run(n, Q.p, n, As.p, n, w, V, ldv, count, err, iters, s);
}
}  // namespace

int zolo_eig_below_ex_device(int precision, int bounds, int rk, int64_t n, const double* A, int64_t lda, double sigma,
                             double* w, double* V, int64_t ldv, int* count, zolo_info_t* info, double* err,
                             int* iters, void* stream) try {
  if (precision < 0 || precision > 4) return -1;
  if (bounds != 0 && bounds != 1) return -2;
  if (rk < 0 || rk > 8) return -3;
  if (n < 1 || n > 32768) return -4;
  if (!A) return -5;
  if (lda < n) return -6;
  if (!std::isfinite(sigma)) return -7;
  if (!w) return -8;
  if (!V) return -9;
  if (ldv < n) return -10;
  if (!count) return -11;
  return eig_below_run(precision, bounds, rk, n, A, lda, sigma, w, V, ldv, count, info, err, iters, stream);
} catch (...) {
  return exception_status();
}

int zolo_eig_below_device(int precision, int rk, int64_t n, const double* A, int64_t lda, double sigma, double* w,
                          double* V, int64_t ldv, int* count, zolo_info_t* info, double* err, int* iters,
                          void* stream) try {
  if (precision < 0 || precision > 4) return -1;
  if (rk < 0 || rk > 8) return -2;
  if (n < 1 || n > 32768) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!std::isfinite(sigma)) return -6;
  if (!w) return -7;
  if (!V) return -8;
  if (ldv < n) return -9;
  if (!count) return -10;
  return eig_below_run(precision, 0, rk, n, A, lda, sigma, w, V, ldv, count, info, err, iters, stream);
} catch (...) {
  return exception_status();
}

int zolo_subit_host(int64_t n, double* A, int64_t lda, double normA, double* Q, int64_t ldq, int* rk, double* err,
                    int* iters) try {
  if (n < 1 || n > 32768) return -1;
  if (!A) return -2;
  if (lda < n) return -3;
  if (!(normA > 0.0)) return -4;
  if (!Q) return -5;
  if (ldq < n) return -6;
  if (!rk) return -7;
  ZOLO_CHECK(ensure_device());
  DevMat dA, dQ;
  ZOLO_CHECK(dA.alloc(n));
  ZOLO_CHECK(dQ.alloc(n));
  ZOLO_CHECK(dA.upload(A, lda, n));
  ZOLO_CHECK(dQ.upload(Q, ldq, n));
  ZOLO_CHECK(subit_run(n, dA.p, n, normA, dQ.p, n, rk, err, iters, nullptr));
  ZOLO_CHECK(dA.download(A, lda, n));
  ZOLO_CHECK(dQ.download(Q, ldq, n));
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_eig_host(int64_t n, double* Q, int64_t ldq, double* As, int64_t lda, double* w, double* V, int64_t ldv,
                  int* rk) try {
  if (n < 1 || n > 32768) return -1;
  if (!Q) return -2;
  if (ldq < n) return -3;
  if (!As) return -4;
  if (lda < n) return -5;
  if (!w) return -6;
  if (!V) return -7;
  if (ldv < n) return -8;
  if (!rk) return -9;
  ZOLO_CHECK(ensure_device());
  DevMat dQ, dA, dV;
  ZOLO_CHECK(dQ.alloc(n));
  ZOLO_CHECK(dA.alloc(n));
  ZOLO_CHECK(dV.alloc(n));
  double* dw = nullptr;
  ZOLO_HIP(hipMalloc(&dw, sizeof(double) * n));
  struct Free {
    double* p;
    ~Free() { (void)hipFree(p); }
  } fw{dw};
  ZOLO_CHECK(dQ.upload(Q, ldq, n));
  ZOLO_CHECK(dA.upload(As, lda, n));
  int k = 0;
  ZOLO_CHECK(zoloeig_run(n, dQ.p, n, dA.p, n, dw, dV.p, n, &k, nullptr, nullptr, nullptr));
  ZOLO_HIP(hipDeviceSynchronize());
  if (k > 0) {
    ZOLO_HIP(hipMemcpy(w, dw, sizeof(double) * k, hipMemcpyDeviceToHost));
    ZOLO_HIP(hipMemcpy2D(V, sizeof(double) * ldv, dV.p, sizeof(double) * n, sizeof(double) * n, k,
                         hipMemcpyDeviceToHost));
  }
  *rk = k;
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_mm_info(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz, int* mtype, int64_t* bandwidth) try {
  return mm_info(path, rows, cols, nnz, mtype, bandwidth);
} catch (...) {
  return exception_status();
}

int zolo_mm_dense_host(const char* path, double sigma, double* A, int64_t lda, int* mtype) try {
  return mm_dense_host(path, sigma, A, lda, mtype);
} catch (...) {
  return exception_status();
}

int zolo_mm_dense_device(const char* path, double sigma, double* A, int64_t lda, int* mtype, void* stream) try {
  ZOLO_CHECK(ensure_device());
  return mm_dense_device(path, sigma, A, lda, mtype, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_formx2_host(int64_t n, double fone, double* A, int64_t lda, double* Z, int64_t ldz, double* err) try {
  if (n < 1) return -1;
  if (!(fone != 0.0)) return -2;
  if (!A) return -3;
  if (lda < n) return -4;
  if (!Z) return -5;
  if (ldz < n) return -6;
  if (!err) return -7;
  ZOLO_CHECK(ensure_device());
  DevMat dA, dZ;
  ZOLO_CHECK(dA.alloc(n));
  ZOLO_CHECK(dZ.alloc(n));
  ZOLO_CHECK(dA.upload(A, lda, n));
  ZOLO_CHECK(dZ.upload(Z, ldz, n));
  constexpr int kParts = 1024;
  const size_t bytes = sizeof(double) * (kParts + 8) + 512;
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  double* part = ar.take<double>(kParts);
  double* out = ar.take<double>(8);
  // X <- (X + Z) / FONE with the error, on the device (FormX2.f90:73-89)
  ZOLO_CHECK(k_formx2<double>(dA.p, n, dZ.p, n, n * n, 1, n, fone, part, kParts, out, nullptr));
  double e2 = 0.0;
  ZOLO_HIP(hipMemcpy(&e2, out, sizeof(double), hipMemcpyDeviceToHost));
  *err = std::sqrt(e2);
  ZOLO_CHECK(dA.download(A, lda, n));
  ZOLO_CHECK(dA.download(Z, ldz, n));  // FormX2 leaves the new iterate in both arrays
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_syevd_host(int64_t n, double* H, int64_t ldh, double* w, double* Z, int64_t ldz) try {
  if (n < 1 || n > 32768) return -1;
  if (!H) return -2;
  if (ldh < n) return -3;
  if (!w) return -4;
  if (!Z) return -5;
  if (ldz < n) return -6;
  ZOLO_CHECK(ensure_device());
  DevMat dH, dZ;
  ZOLO_CHECK(dH.alloc(n));
  ZOLO_CHECK(dZ.alloc(n));
  double* dw = nullptr;
  ZOLO_HIP(hipMalloc(&dw, sizeof(double) * n));
  struct Free {
    double* p;
    ~Free() { (void)hipFree(p); }
  } fw{dw};
  ZOLO_CHECK(dH.upload(H, ldh, n));
  ZOLO_CHECK(syevd_run(n, dH.p, n, dw, dZ.p, n, nullptr));
  ZOLO_HIP(hipDeviceSynchronize());
  ZOLO_HIP(hipMemcpy(w, dw, sizeof(double) * n, hipMemcpyDeviceToHost));
  ZOLO_CHECK(dZ.download(Z, ldz, n));
  return kOk;
} catch (...) {
  return exception_status();
}

namespace {
// ZoloCHK1.f90:80-91 on validated arguments (bounds: 0 = from the first QR, 1 = LU)
int svd_run(int precision, int bounds, int rk, int64_t n, const double* A, int64_t lda, double* S, double* U,
            int64_t ldu, double* V, int64_t ldv, zolo_info_t* info, double* phase_ms, void* stream) {
  ZOLO_CHECK(ensure_device());
  hipStream_t s = (hipStream_t)stream;
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  DevMat Ht;  // H, then U_p V
  ZOLO_CHECK(Ht.alloc(n));
  zolo_info_t local{};
  zolo_info_t* inf = info ? info : &local;
  // bounds = 1: alpha, beta by LU + 1-norm estimate as the reference's driver
  // computes them (test_zolodrv1.f90:179-189); 0: from the QR the first step
  // reuses (alpha = beta = 0 below)
  double alpha = 0.0, beta = 0.0;
  if (bounds == 1) {
    ZOLO_CHECK(zolo_bounds_device(0, n, A, lda, &alpha, &beta, stream));
    if (rk == 0) rk = choosem(1.0 / beta);  // the driver's rule (pzolopd1.f90 is handed r)
  }
  // polar factor U_p into U, H = (U_p^T A + A^T U_p)/2 into Ht (ZoloCHK1.f90:80-83)
  // (rk = 0: choosem)
  const int st = zolo_pd_run(precision, 'N', rk, n, alpha, beta, A, lda, U, ldu, Ht.p, n, inf, s);
  if (st) return st;
  ZOLO_HIP(hipStreamSynchronize(s));
  const auto t1 = clk::now();
  // H = V diag(S) V^T (ZoloCHK1.f90:86), S ascending as ELPA returns it
  double ph[3] = {0, 0, 0};
  ZOLO_CHECK(syevd_run(n, Ht.p, n, S, V, ldv, s, ph));
  // U = U_p V (ZoloCHK1.f90:90)
  GemmDesc g;
  g.M = n; g.N = n; g.K = n;
  g.A = U; g.lda = ldu; g.B = V; g.ldb = ldv; g.C = Ht.p; g.ldc = n;
  ZOLO_CHECK(gemm(Prec::F64, g, s));
  ZOLO_HIP(hipMemcpy2DAsync(U, sizeof(double) * ldu, Ht.p, sizeof(double) * n, sizeof(double) * n, n,
                            hipMemcpyDeviceToDevice, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  const auto t2 = clk::now();
  if (phase_ms) {
    phase_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();  // bounds + polar + H
    phase_ms[1] = ph[0];                                                           // tridiagonalisation
    phase_ms[2] = ph[1];                                                           // divide and conquer
    phase_ms[3] = ph[2];                                                           // back-transformation
    phase_ms[4] = std::chrono::duration<double, std::milli>(t2 - t1).count();   // whole eigensolve + U
  }
  return kOk;
}
}  // namespace

int zolo_svd_ex_device(int precision, int bounds, int rk, int64_t n, const double* A, int64_t lda, double* S,
                       double* U, int64_t ldu, double* V, int64_t ldv, zolo_info_t* info, double* phase_ms,
                       void* stream) try {
  if (precision < 0 || precision > 4) return -1;
  if (bounds != 0 && bounds != 1) return -2;
  if (rk < 0 || rk > 8) return -3;
  if (n < 1 || n > 32768) return -4;
  if (!A) return -5;
  if (lda < n) return -6;
  if (!S) return -7;
  if (!U) return -8;
  if (ldu < n) return -9;
  if (!V) return -10;
  if (ldv < n) return -11;
  return svd_run(precision, bounds, rk, n, A, lda, S, U, ldu, V, ldv, info, phase_ms, stream);
} catch (...) {
  return exception_status();
}

int zolo_svd_device(int precision, int rk, int64_t n, const double* A, int64_t lda, double* S, double* U, int64_t ldu,
                    double* V, int64_t ldv, zolo_info_t* info, double* phase_ms, void* stream) try {
  if (precision < 0 || precision > 4) return -1;
  if (rk < 0 || rk > 8) return -2;
  if (n < 1 || n > 32768) return -3;
  if (!A) return -4;
  if (lda < n) return -5;
  if (!S) return -6;
  if (!U) return -7;
  if (ldu < n) return -8;
  if (!V) return -9;
  if (ldv < n) return -10;
  return svd_run(precision, 0, rk, n, A, lda, S, U, ldu, V, ldv, info, phase_ms, stream);
} catch (...) {
  return exception_status();
}

int zolo_l0_param_count(void) { return kParamCount; }

int zolo_l0_set_weights(const float* params, int nparams) try {
  if (!params) return -1;
  if (nparams != kParamCount) {
    set_error("zolo_l0_set_weights: expected " + std::to_string(kParamCount) + " parameters");
    return -2;
  }
  for (int i = 0; i < kFeat; ++i)
    if (!(params[kFeat + i] > 0.0f)) {
      set_error("zolo_l0_set_weights: feature scales must be positive");
      return -1;
    }
  ZOLO_CHECK(ensure_device());
  return l0_set_weights(params);
} catch (...) {
  return exception_status();
}

int zolo_l0_features_device(int count, const int64_t* n, const float* const* A, const int64_t* lda, float* feat,
                            float* smax, void* stream) try {
  if (count < 0) return -1;
  if (count > 0 && (!n || !A || !lda || !feat || !smax)) return -2;
  for (int i = 0; i < count; ++i) {
    if (n[i] < 1 || n[i] > kMaxBatchN) return -2;
    if (lda[i] < n[i]) return -4;
  }
  if (count == 0) return kOk;
  ZOLO_CHECK(ensure_device());
  return l0_features(count, n, A, lda, feat, smax, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_l0_forward_device(const float* params, const float* feat, int N, float* y, void* stream) try {
  if (!params) return -1;
  if (!feat) return -2;
  if (N < 0) return -3;
  if (!y) return -4;
  ZOLO_CHECK(ensure_device());
  ZOLO_CHECK(kl_forward(params, feat, N, nullptr, nullptr, y, (hipStream_t)stream));
  ZOLO_HIP(hipStreamSynchronize((hipStream_t)stream));
  return kOk;
} catch (...) {
  return exception_status();
}

int zolo_l0_grad_device(const float* params, const float* feat, const float* target, int N, float tau, float* grad,
                        float* loss, void* stream) try {
  if (!params) return -1;
  if (!feat) return -2;
  if (!target) return -3;
  if (N < 1) return -4;
  if (!(tau > 0.0f && tau < 1.0f)) return -5;
  if (!grad) return -6;
  ZOLO_CHECK(ensure_device());
  return l0_grad(params, feat, target, N, tau, grad, loss, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

int zolo_l0_train_device(float* params, const float* feat, const float* target, int N, int steps, float lr, float tau,
                         float* loss_hist, void* stream) try {
  if (!params) return -1;
  if (!feat) return -2;
  if (!target) return -3;
  if (N < 1) return -4;
  if (steps < 0) return -5;
  if (!(lr > 0.0f)) return -6;
  if (!(tau > 0.0f && tau < 1.0f)) return -7;
  ZOLO_CHECK(ensure_device());
  return l0_train(params, feat, target, N, steps, lr, tau, loss_hist, (hipStream_t)stream);
} catch (...) {
  return exception_status();
}

}  // extern "C"
