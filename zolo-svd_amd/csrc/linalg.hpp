// Blocked dense factorizations built on the MFMA GEMM engine.
#pragma once
#include <string>
#include <vector>

#include "kernels.hpp"

namespace zolo {

// Device scratch owned by the library: one grow-only allocation per process,
// carved by a bump allocator for each top-level call.
class Arena {
 public:
  explicit Arena(char* base = nullptr, size_t cap = 0) : base_(base), cap_(cap), off_(0) {}
  template <typename T>
  T* take(size_t count) {
    off_ = (off_ + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base_ + off_);
    off_ += count * sizeof(T);
    return p;
  }
  size_t used() const { return off_; }
  bool overflow() const { return off_ > cap_; }
  void reset() { off_ = 0; }
 private:
  char* base_;
  size_t cap_;
  size_t off_;
};

#define ZOLO_ARENA_OK(ar)                                             \
  do {                                                                \
    if ((ar).overflow()) {                                            \
      ::zolo::set_error(std::string("workspace overflow in ") + __func__); \
      return ::zolo::kErrInternal;                                    \
    }                                                                 \
  } while (0)

// Grow-only device workspace shared by the entry points (not re-entrant).
int workspace_acquire(size_t bytes, char** base);

// Products of the mixed precisions' FP64 parts (zolo_set_f64_products): 1 =
// exact int8 products (default), 0 = FP64 MFMA.
int f64_products();
void set_f64_products(int mode);
void workspace_release_all();

struct Ctx {
  hipStream_t s = nullptr;
  Prec low = Prec::F32;  // GEMM precision used for float operands (F32 or BF16X3)
  // GEMM precision of the large square FP64 products (the Zolotarev sum
  // sum_q alpha_q Q1_q Q2_q^T, the polish's U^T U and U^T X, H = U^T A):
  // F64, or F64X (exact int8 products, FP64-level accuracy) in the mixed precisions
  Prec big = Prec::F64;
  int tile = 0;          // GEMM tile hint for products with M, N > 64 (0: automatic)
};

// Applies the context's tile hint to a GEMM descriptor.
inline GemmDesc with_tile(const Ctx& c, GemmDesc d) {
  if (c.tile && d.tile == 0 && d.M > 64 && d.N > 64) d.tile = c.tile;
  return d;
}

template <typename T> inline Prec prec_for(const Ctx& c) { return Prec::F32; }
template <> inline Prec prec_for<double>(const Ctx&) { return Prec::F64; }
template <> inline Prec prec_for<float>(const Ctx& c) { return c.low; }
// precision of the large square products: the context's choice for FP64
template <typename T> inline Prec big_prec(const Ctx& c) { return prec_for<T>(c); }
template <> inline Prec big_prec<double>(const Ctx& c) { return c.big; }

// Batched Cholesky with explicit inverse: A_q (lower) -> L_q, Li_q = L_q^{-1}.
// tmp must hold batch * ceil(n/2)^2 elements (stride sT).
template <typename T>
int chol_inv(const Ctx& c, T* A, int64_t lda, int64_t sA, T* Li, int64_t ldl, int64_t sL, int64_t n, T* tmp,
             int64_t sT, int* info, int batch);

// Z_q = alpha_q * X (scale X^T X + (shift_q + add) I)^{-1}, q = 0..r-1 (one batch);
// left = true gives alpha_q * (scale X^T X + (shift_q + add) I)^{-1} X, the
// SYM='S' branch of zolochol.f90:99-100 / pqdwhfacs.f90:425.
template <typename T>
size_t chol_step_bytes(int64_t n, int r);
template <typename T>
int chol_step(const Ctx& c, const T* X, int64_t ldx, int64_t n, int r, const double* d_shift, double scale, double add,
              const double* d_alpha, T* Z, int64_t ldz, int64_t sZ, Arena& ar, int* d_info, bool left = false);

// Z_q = alpha_q * Q1_q Q2_q^T where [X; s_q I] = [Q1_q; Q2_q] R_q (structured
// Householder QR).  With r >= 2 shifts on a large X the step factors X = Q_X R
// once and takes the QR of [R; s_q I] per shift: Q1_q = Q_X Q~1_q, Q2_q =
// Q~2_q with both Q~ blocks upper triangular, so Z = Q_X sum_q alpha_q Q~1_q
// Q~2_q^T costs about 2 n^3 per shift instead of 5 n^3.  The summed Z is then
// returned in slot 0 and *nz = 1; otherwise the r slots hold Z_q and *nz = r.
template <typename T>
size_t qr_step_bytes(int64_t n, int r);
// A QR of the iterate computed beforehand for the factored form: Xw holds R
// (upper triangle, already scaled to the iterate) and the reflectors below it,
// Tx the outer T factors (dense_geqrf's storage).
template <typename T>
struct QrPre {
  T* Xw;
  T* Tx;
};
template <typename T>
int qr_step(const Ctx& c, const T* X, int64_t ldx, int64_t n, int r, const double* d_sdiag, const double* d_alpha,
            T* Z, int64_t ldz, int64_t sZ, Arena& ar, int* nz, const QrPre<T>* pre = nullptr);

// Row structure of a blocked Householder QR: Stack = [X; s I] (mpdgeqrf.f),
// Dense = one n x n matrix, Tri = [R; s I] with R upper triangular.
enum class QrShape { Stack, Dense, Tri };

template <typename T>
size_t stacked_qr_bytes(int64_t n, int r);

// Structured GEQRF (R + reflectors in place, tau out) and ORGQR (explicit
// 2n x n Q from the reflectors) of a batch of 2n x n stacks; stacked_qr does both.
template <typename T>
int stacked_geqrf(const Ctx& c, T* S, int64_t lds, int64_t sS, int64_t n, int batch, T* tau, Arena& ar);
template <typename T>
int stacked_orgqr(const Ctx& c, const T* S, int64_t lds, int64_t sS, const T* tau, T* Q, int64_t ldq, int64_t sQ,
                  int64_t n, int batch, Arena& ar);
template <typename T>
int stacked_qr(const Ctx& c, T* S, int64_t lds, int64_t sS, T* Q, int64_t ldq, int64_t sQ, int64_t n, int batch,
               T* tau_out, Arena& ar, QrShape sh = QrShape::Stack);

// Dense n x n Householder QR of a batch in place (tau: batch x n, Tall: batch x
// ceil(n/256) x 256^2 outer T factors) and M <- Q M with the reflectors it left.
template <typename T>
size_t dense_qr_bytes(int64_t n, int batch = 1);
template <typename T>
int dense_geqrf(const Ctx& c, T* A, int64_t lda, int64_t sA, int64_t n, int batch, T* tau, T* Tall, Arena& ar);
template <typename T>
int dense_apply_q(const Ctx& c, const T* A, int64_t lda, int64_t sA, int64_t n, int batch, const T* Tall, T* M,
                  int64_t ldm, int64_t sM, Arena& ar);
// Tall n x k Householder QR (k <= n; the same storage as dense_geqrf, bytes
// dense_qr_bytes(n)) and Q (n x m) = the first m columns of its orthogonal
// factor (m <= n; PDORGQR(N, M, K)).
template <typename T>
int dense_geqrf_cols(const Ctx& c, T* A, int64_t lda, int64_t n, int64_t k, T* tau, T* Tall, Arena& ar);
template <typename T>
int dense_orgqr_cols(const Ctx& c, const T* A, int64_t lda, int64_t n, int64_t k, const T* Tall, T* Q, int64_t ldq,
                     int64_t m, Arena& ar);

}  // namespace zolo
