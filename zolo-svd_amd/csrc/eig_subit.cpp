// Spectral division of a symmetric matrix by its polar factor: the invariant
// subspace of the eigenvalues below a shift sigma and their eigenpairs.
//
//  * subit_run   replaces PQDWHSUBIT (pqdwhsubit.f90:2-316): from the polar
//    factor Q of A - sigma I, C = (Q - I)/2 is minus the spectral projector
//    onto the eigenvalues below sigma, Rk = round(||C||_F^2) their number.  A
//    random N x (Rk + 3) block is orthonormalised, multiplied by C and
//    orthonormalised again with the full N x N Q of that QR, V = [V1 V2];
//    E = V^T A V is accepted when ||E(1:Rk, Rk+1:N)||_F / ||A||_F <= 4 TOL1
//    (TOL1 = 10 eps / 2), otherwise one more iteration starts from C V1
//    (at most two, as the reference).  On exit Q holds V and A holds E.
//  * zoloeig_run replaces ZoloEig (ZoloEig.f90:2-110): PQDWHSUBIT, then the
//    symmetric eigensolver on the leading Rk x Rk block of E and the
//    eigenvectors V1 VV1.
//
// All products are the library's FP64 MFMA GEMMs, the QR factorisations the
// blocked Householder GEQRF of the n x kk block with its block reflectors
// applied to [I; 0] (PDORGQR(N, KK, KK) and PDORGQR(N, N, KK)), the
// eigensolver syevd_run.
#include <algorithm>
#include <cmath>

#include "coeff.hpp"
#include "driver.hpp"
#include "eig.hpp"
#include "kernels.hpp"

namespace zolo {
namespace {

// Device buffer owned by one call (the shared workspace is used by the QR and
// the eigensolver in between, so nothing here lives in it across calls).
struct Buf {
  double* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (hipMalloc(&p, sizeof(double) * std::max<size_t>(count, 1)) != hipSuccess) {
      p = nullptr;
      set_error("spectral division: hipMalloc failed");
      return kErrAlloc;
    }
    return kOk;
  }
};

// Squared Frobenius norm of an m x n block (0 for an empty block).
int fro2(const double* A, int64_t lda, int64_t m, int64_t n, double* part, double* out, double* h, hipStream_t s) {
  *h = 0.0;
  if (m <= 0 || n <= 0) return kOk;
  ZOLO_CHECK(k_sumsq<double>(A, lda, m, n, part, 256, out, s));
  ZOLO_HIP(hipMemcpyAsync(h, out, sizeof(double), hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  return kOk;
}

// Q (n x m) = the first m columns of the orthogonal factor of the QR of the
// n x k block X (overwritten by the reflectors): GEQRF of k columns, then
// their block reflectors applied to [I; 0] (PDGEQRF + PDORGQR(N, M, K)).
int qr_q(const Ctx& c, double* X, int64_t n, int64_t k, double* Q, int64_t m) {
  const int64_t nout = (n + 255) / 256;
  const size_t bytes = dense_qr_bytes<double>(n) + sizeof(double) * ((size_t)n + (size_t)nout * 256 * 256) + 8192;
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  double* tau = ar.take<double>((size_t)n);
  double* Tall = ar.take<double>((size_t)nout * 256 * 256);
  ZOLO_ARENA_OK(ar);
  const size_t mark = ar.used();
  ZOLO_CHECK(dense_geqrf_cols<double>(c, X, n, n, k, tau, Tall, ar));
  // the factorisation's scratch is free again for forming Q
  ar.reset();
  ar.take<char>(mark);
  ZOLO_CHECK(dense_orgqr_cols<double>(c, X, n, n, k, Tall, Q, n, m, ar));
  return kOk;
}

// C (m x k) = A (m x kk) B (kk x k), all with explicit leading dimensions.
int gemm_nn(const Ctx& c, int ta, int64_t m, int64_t k, int64_t kk, const double* A, int64_t lda, const double* B,
            int64_t ldb, double* C, int64_t ldc) {
  if (m <= 0 || k <= 0 || kk <= 0) return kOk;
  GemmDesc d;
  d.transA = ta; d.M = m; d.N = k; d.K = kk;
  d.A = A; d.lda = lda; d.B = B; d.ldb = ldb; d.C = C; d.ldc = ldc;
  return gemm(Prec::F64, d, c.s);
}

}  // namespace

int subit_run(int64_t n, double* A, int64_t lda, double normA, double* Q, int64_t ldq, int* rk_out, double* err_out,
              int* it_out, hipStream_t s) {
  Ctx c;
  c.s = s;
  const int64_t nn = n * n;
  Buf X, V, W, red;
  ZOLO_CHECK(X.alloc((size_t)nn));
  ZOLO_CHECK(V.alloc((size_t)nn));
  ZOLO_CHECK(W.alloc((size_t)nn));
  ZOLO_CHECK(red.alloc(512));
  double h = 0.0;
  // C = (Q - I) / 2 in place (pqdwhsubit.f90, the diagonal shift and PDLASCL)
  ZOLO_CHECK(k_scale_copy<double, double>(Q, ldq, Q, ldq, n, n, 0.5, s));
  ZOLO_CHECK(k_add_diag<double>(Q, ldq, n, -0.5, s));
  ZOLO_CHECK(fro2(Q, ldq, n, n, red.p, red.p + 256, &h, s));
  const int64_t rk = std::min<int64_t>(n, std::max<int64_t>(0, (int64_t)std::llround(h)));
  int64_t kk = std::min<int64_t>(rk + 3, n);  // oversampling 3
  const double tol1 = 10.0 * dlamch_p() / 2.0;
  // random n x kk start block
  ZOLO_CHECK(k_fill_uniform(X.p, n, n, kk, 0x5eedull + (unsigned long long)n, s));
  double err = 0.0;
  int it = 1;
  for (;;) {
    // X1 (n x kk) = orthonormal basis of X (pqdwhsubit.f90:266-267)
    ZOLO_CHECK(qr_q(c, X.p, n, kk, V.p, kk));
    // Y = C X1 (:269), then V = [V1 V2] (n x n) from the QR of Y (:271-272)
    ZOLO_CHECK(gemm_nn(c, 0, n, kk, n, Q, ldq, V.p, n, X.p, n));
    ZOLO_CHECK(qr_q(c, X.p, n, kk, V.p, n));
    // E = V^T A V into X
    ZOLO_CHECK(gemm_nn(c, 0, n, n, n, A, lda, V.p, n, W.p, n));
    ZOLO_CHECK(gemm_nn(c, 1, n, n, n, V.p, n, W.p, n, X.p, n));
    ZOLO_CHECK(fro2(X.p + rk * n, n, rk, n - rk, red.p, red.p + 256, &h, s));
    err = std::sqrt(h) / normA;
    if (err <= 4.0 * tol1 || it == 2 || rk == 0) break;
    // one more iteration from X = C V1 (kk = Rk, :305-311)
    ZOLO_CHECK(gemm_nn(c, 0, n, rk, n, Q, ldq, V.p, n, W.p, n));
    std::swap(X.p, W.p);
    kk = rk;
    ++it;
  }
  // Q <- V, A <- E
  ZOLO_HIP(hipMemcpy2DAsync(Q, sizeof(double) * ldq, V.p, sizeof(double) * n, sizeof(double) * n, n,
                            hipMemcpyDeviceToDevice, s));
  ZOLO_HIP(hipMemcpy2DAsync(A, sizeof(double) * lda, X.p, sizeof(double) * n, sizeof(double) * n, n,
                            hipMemcpyDeviceToDevice, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  *rk_out = (int)rk;
  if (err_out) *err_out = err;
  if (it_out) *it_out = it;
  return kOk;
}

int zoloeig_run(int64_t n, double* Q, int64_t ldq, double* As, int64_t lda, double* w, double* Vout, int64_t ldv,
                int* rk_out, double* err_out, int* it_out, hipStream_t s) {
  Ctx c;
  c.s = s;
  Buf red;
  ZOLO_CHECK(red.alloc(512));
  double h = 0.0;
  ZOLO_CHECK(fro2(As, lda, n, n, red.p, red.p + 256, &h, s));
  const double normA = std::sqrt(h);
  if (!(normA > 0.0)) {
    set_error("spectral division: zero matrix");
    return kErrArg;
  }
  int rk = 0;
  ZOLO_CHECK(subit_run(n, As, lda, normA, Q, ldq, &rk, err_out, it_out, s));
  *rk_out = rk;
  if (rk == 0) return kOk;
  // eigenpairs of the leading Rk x Rk block of E (ZoloEig.f90: PDSYEVD on Rk)
  Buf Ek, Z;
  ZOLO_CHECK(Ek.alloc((size_t)rk * rk));
  ZOLO_CHECK(Z.alloc((size_t)rk * rk));
  ZOLO_HIP(hipMemcpy2DAsync(Ek.p, sizeof(double) * rk, As, sizeof(double) * lda, sizeof(double) * rk, rk,
                            hipMemcpyDeviceToDevice, s));
  ZOLO_CHECK(syevd_run(rk, Ek.p, rk, w, Z.p, rk, s));
  // eigenvectors of A: V1 VV1 (n x Rk)
  ZOLO_CHECK(gemm_nn(c, 0, n, rk, rk, Q, ldq, Z.p, rk, Vout, ldv));
  ZOLO_HIP(hipStreamSynchronize(s));
  return kOk;
}

}  // namespace zolo
