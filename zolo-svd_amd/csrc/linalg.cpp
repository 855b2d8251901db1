// Blocked factorizations on the MFMA GEMM engine: recursive Cholesky with an
// explicit triangular inverse, the Zolotarev/QDWH shifted Cholesky step
// (zolochol.f90:87-113, pqdwhfacs.f90:404-441) and the structured Householder
// QR step (zoloqr.f90:82-112 over mpdgeqrf.f / mpdorgqr.f).
//
// Shifted solves use Z = X L^{-T} L^{-1} with L^{-1} formed once per shift:
// two triangular GEMMs replace the transpose + PDPOSV + transpose sequence of
// the reference and keep every O(n^3) operation on the MFMA GEMM engine.
#include <algorithm>
#include <cstring>

#include "linalg.hpp"

namespace zolo {

namespace {
char* g_ws = nullptr;
size_t g_ws_cap = 0;

template <typename T>
int gemm_t(const Ctx& c, const GemmDesc& d) {
  return gemm(prec_for<T>(c), with_tile(c, d), c.s);
}
}  // namespace

int workspace_acquire(size_t bytes, char** base) {
  if (bytes > g_ws_cap) {
    if (g_ws) {
      ZOLO_HIP(hipDeviceSynchronize());
      ZOLO_HIP(hipFree(g_ws));
      g_ws = nullptr;
      g_ws_cap = 0;
    }
    hipError_t e = hipMalloc(&g_ws, bytes);
    if (e != hipSuccess) {
      g_ws = nullptr;
      set_error("workspace: hipMalloc of " + std::to_string(bytes) + " bytes failed");
      return kErrAlloc;
    }
    g_ws_cap = bytes;
  }
  *base = g_ws;
  return kOk;
}

void workspace_release_all() {
  if (g_ws) {
    (void)hipDeviceSynchronize();
    (void)hipFree(g_ws);
  }
  g_ws = nullptr;
  g_ws_cap = 0;
}

// ---------------------------------------------------------------------------
// Recursive Cholesky + inverse.
// ---------------------------------------------------------------------------
template <typename T>
static int chol_inv_rec(const Ctx& c, T* A, int64_t lda, int64_t sA, T* Li, int64_t ldl, int64_t sL, int64_t n,
                        int64_t col0, T* tmp, int64_t sT, int* info, int batch) {
  if (n <= kCholLeaf) return k_chol_leaf<T>(A, lda, sA, Li, ldl, sL, (int)n, col0, info, batch, c.s);
  int64_t n1 = ((n / 2 + kCholLeaf - 1) / kCholLeaf) * kCholLeaf;
  if (n1 >= n) n1 = n - kCholLeaf;
  const int64_t n2 = n - n1;
  T* A11 = A;
  T* A21 = A + n1;
  T* A22 = A + n1 + n1 * lda;
  T* L11 = Li;
  T* L21 = Li + n1;
  T* L22 = Li + n1 + n1 * ldl;
  ZOLO_CHECK(chol_inv_rec<T>(c, A11, lda, sA, L11, ldl, sL, n1, col0, tmp, sT, info, batch));
  GemmDesc d;
  d.batch = batch;
  // tmp = A21 * Li11^T  (= L21), op(B) upper in (k,j)
  d.transA = 0; d.transB = 1; d.M = n2; d.N = n1; d.K = n1;
  d.A = A21; d.lda = lda; d.strideA = sA;
  d.B = L11; d.ldb = ldl; d.strideB = sL;
  d.C = tmp; d.ldc = n2; d.strideC = sT;
  d.alpha = 1.0; d.beta = 0.0; d.bHi = 0;
  ZOLO_CHECK(gemm_t<T>(c, d));
  // A22 -= L21 L21^T (lower triangle)
  d = GemmDesc(); d.batch = batch;
  d.transA = 0; d.transB = 1; d.M = n2; d.N = n2; d.K = n1;
  d.A = tmp; d.lda = n2; d.strideA = sT;
  d.B = tmp; d.ldb = n2; d.strideB = sT;
  d.C = A22; d.ldc = lda; d.strideC = sA;
  d.alpha = -1.0; d.beta = 1.0; d.lowerC = 1;
  ZOLO_CHECK(gemm_t<T>(c, d));
  // A21 <- L21 * Li11  (first half of Li21), op(B) lower in (k,j)
  d = GemmDesc(); d.batch = batch;
  d.transA = 0; d.transB = 0; d.M = n2; d.N = n1; d.K = n1;
  d.A = tmp; d.lda = n2; d.strideA = sT;
  d.B = L11; d.ldb = ldl; d.strideB = sL;
  d.C = A21; d.ldc = lda; d.strideC = sA;
  d.alpha = 1.0; d.beta = 0.0; d.bLo = 0;
  ZOLO_CHECK(gemm_t<T>(c, d));
  ZOLO_CHECK(chol_inv_rec<T>(c, A22, lda, sA, L22, ldl, sL, n2, col0 + n1, tmp, sT, info, batch));
  // Li21 = -Li22 * (L21 Li11), op(A) lower
  d = GemmDesc(); d.batch = batch;
  d.transA = 0; d.transB = 0; d.M = n2; d.N = n1; d.K = n2;
  d.A = L22; d.lda = ldl; d.strideA = sL;
  d.B = A21; d.ldb = lda; d.strideB = sA;
  d.C = L21; d.ldc = ldl; d.strideC = sL;
  d.alpha = -1.0; d.beta = 0.0; d.aHi = 0;
  ZOLO_CHECK(gemm_t<T>(c, d));
  return kOk;
}

template <typename T>
int chol_inv(const Ctx& c, T* A, int64_t lda, int64_t sA, T* Li, int64_t ldl, int64_t sL, int64_t n, T* tmp,
             int64_t sT, int* info, int batch) {
  // The recursion never writes the strict upper triangle of Li, while its
  // triangular GEMMs (and the ones that follow) trim the K range only at tile
  // granularity and so read part of it: it must be exactly zero beforehand.
  ZOLO_CHECK(k_zero_upper<T>(Li, ldl, sL, n, batch, c.s));
  return chol_inv_rec<T>(c, A, lda, sA, Li, ldl, sL, n, 0, tmp, sT, info, batch);
}

// ---------------------------------------------------------------------------
// Shifted Cholesky step.
// ---------------------------------------------------------------------------
template <typename T>
size_t chol_step_bytes(int64_t n, int r) {
  const int64_t h = (n + 1) / 2 + kCholLeaf;
  return sizeof(T) * (size_t)(n * n + 2 * (size_t)r * n * n + (size_t)r * h * h) + 4096 * 4;
}

template <typename T>
int chol_step(const Ctx& c, const T* X, int64_t ldx, int64_t n, int r, const double* d_shift, double scale, double add,
              const double* d_alpha, T* Z, int64_t ldz, int64_t sZ, Arena& ar, int* d_info, bool left) {
  const int64_t nn = n * n;
  const int64_t h = (n + 1) / 2 + kCholLeaf;
  T* W = ar.take<T>(nn);
  T* Wq = ar.take<T>((size_t)r * nn);
  T* Lq = ar.take<T>((size_t)r * nn);
  T* tmp = ar.take<T>((size_t)r * h * h);
  ZOLO_ARENA_OK(ar);
  // W = X^T X (lower triangle), computed once for all shifts.
  GemmDesc d;
  d.transA = 1; d.transB = 0; d.M = n; d.N = n; d.K = n;
  d.A = X; d.lda = ldx; d.B = X; d.ldb = ldx; d.C = W; d.ldc = n;
  d.alpha = 1.0; d.beta = 0.0; d.lowerC = 1;
  ZOLO_CHECK(gemm_t<T>(c, d));
  // W_q = scale*W + (shift_q + add) I
  ZOLO_CHECK(k_shift_copy<T>(W, n, Wq, n, nn, n, d_shift, scale, add, r, c.s));
  ZOLO_CHECK(chol_inv<T>(c, Wq, n, nn, Lq, n, nn, n, tmp, h * h, d_info, r));
  if (!left) {
    // T_q = X * Li_q^T   (into Wq, which now only holds L_q)
    d = GemmDesc(); d.batch = r;
    d.transA = 0; d.transB = 1; d.M = n; d.N = n; d.K = n;
    d.A = X; d.lda = ldx; d.strideA = 0;
    d.B = Lq; d.ldb = n; d.strideB = nn;
    d.C = Wq; d.ldc = n; d.strideC = nn;
    d.alpha = 1.0; d.beta = 0.0; d.bHi = 0;
    ZOLO_CHECK(gemm_t<T>(c, d));
    // Z_q = alpha_q * T_q * Li_q
    d = GemmDesc(); d.batch = r;
    d.transA = 0; d.transB = 0; d.M = n; d.N = n; d.K = n;
    d.A = Wq; d.lda = n; d.strideA = nn;
    d.B = Lq; d.ldb = n; d.strideB = nn;
    d.C = Z; d.ldc = ldz; d.strideC = sZ;
    d.alpha = 1.0; d.alphaVec = d_alpha; d.beta = 0.0; d.bLo = 0;
    ZOLO_CHECK(gemm_t<T>(c, d));
  } else {
    // T_q = Li_q X ; Z_q = alpha_q Li_q^T T_q
    d = GemmDesc(); d.batch = r;
    d.transA = 0; d.transB = 0; d.M = n; d.N = n; d.K = n;
    d.A = Lq; d.lda = n; d.strideA = nn;
    d.B = X; d.ldb = ldx; d.strideB = 0;
    d.C = Wq; d.ldc = n; d.strideC = nn;
    d.alpha = 1.0; d.beta = 0.0; d.aHi = 0;
    ZOLO_CHECK(gemm_t<T>(c, d));
    d = GemmDesc(); d.batch = r;
    d.transA = 1; d.transB = 0; d.M = n; d.N = n; d.K = n;
    d.A = Lq; d.lda = n; d.strideA = nn;
    d.B = Wq; d.ldb = n; d.strideB = nn;
    d.C = Z; d.ldc = ldz; d.strideC = sZ;
    d.alpha = 1.0; d.alphaVec = d_alpha; d.beta = 0.0; d.aLo = 0;
    ZOLO_CHECK(gemm_t<T>(c, d));
  }
  return kOk;
}

// ---------------------------------------------------------------------------
// Structured Householder QR of [X; s I] and the QR-based step.
// ---------------------------------------------------------------------------
// QR-based step (the structured QR itself is in qr_blocked.cpp).
// ---------------------------------------------------------------------------
namespace {
// The factored form pays off with several shifts; small stacks keep the
// one-workgroup panels of the plain stacked QR.
template <typename T>
bool qr_factored(int64_t n, int r) {
  return r >= 2 && n + kQrPanelWgCols<T> > kQrPanelWgRows;
}
}  // namespace

template <typename T>
size_t qr_step_bytes(int64_t n, int r) {
  size_t b = sizeof(T) * (size_t)4 * r * n * n + stacked_qr_bytes<T>(n, r) + 4096;
  if (qr_factored<T>(n, r)) b += sizeof(T) * (size_t)n * n + dense_qr_bytes<T>(n) + 4096;
  return b;
}

template <typename T>
int qr_step(const Ctx& c, const T* X, int64_t ldx, int64_t n, int r, const double* d_sdiag, const double* d_alpha,
            T* Z, int64_t ldz, int64_t sZ, Arena& ar, int* nz, const QrPre<T>* pre) {
  const int64_t ls = 2 * n, ss = 2 * n * n;
  T* S = ar.take<T>((size_t)r * ss);
  T* Q = ar.take<T>((size_t)r * ss);
  ZOLO_ARENA_OK(ar);
  if (!qr_factored<T>(n, r)) {
    ZOLO_CHECK(k_build_stack<T>(X, ldx, S, ls, ss, n, d_sdiag, r, c.s));
    ZOLO_CHECK(stacked_qr<T>(c, S, ls, ss, Q, ls, ss, n, r, nullptr, ar));
    // Z_q = alpha_q * Q1 Q2^T ; Q2 is exactly zero below its 255th subdiagonal
    // (the block reflectors applied to Q are 256 columns wide).
    GemmDesc d; d.batch = r;
    d.transA = 0; d.transB = 1; d.M = n; d.N = n; d.K = n;
    d.A = Q; d.lda = ls; d.strideA = ss;
    d.B = Q + n; d.ldb = ls; d.strideB = ss;
    d.C = Z; d.ldc = ldz; d.strideC = sZ;
    d.alpha = 1.0; d.alphaVec = d_alpha; d.beta = 0.0; d.bLo = 256;
    ZOLO_CHECK(gemm(big_prec<T>(c), with_tile(c, d), c.s));
    *nz = r;
    return kOk;
  }
  // X = Q_X R (X itself is kept: it is the iterate FormX2 needs), or the
  // caller's factorisation of it
  const int64_t nout = (n + 255) / 256;
  T* Xw = nullptr;
  T* Tx = nullptr;
  if (pre) {
    Xw = pre->Xw;
    Tx = pre->Tx;
  } else {
    Xw = ar.take<T>((size_t)n * n);
    T* taux = ar.take<T>((size_t)n);
    Tx = ar.take<T>((size_t)nout * 256 * 256);
    ZOLO_ARENA_OK(ar);
    ZOLO_CHECK(k_scale_copy<T, T>(X, ldx, Xw, n, n, n, 1.0, c.s));
    ZOLO_CHECK(dense_geqrf<T>(c, Xw, n, 0, n, 1, taux, Tx, ar));
  }
  // [R; s_q I] = [Q~1_q; Q~2_q] R~_q for every shift; rows sorted large to small
  // (R on top) keep the Householder QR row-wise accurate for tiny s_q
  ZOLO_CHECK(k_build_stack<T>(Xw, n, S, ls, ss, n, d_sdiag, r, c.s, true));
  const size_t mark = ar.used();
  ZOLO_CHECK(stacked_qr<T>(c, S, ls, ss, Q, ls, ss, n, r, nullptr, ar, QrShape::Tri));
  ar.reset();
  ar.take<char>(mark);
  // Z~ = sum_q alpha_q Q~1_q Q~2_q^T: both factors are zero below their 255th
  // subdiagonal (256-wide outer blocks), so only k >= max(i, j) - 255 counts
  for (int q = 0; q < r; ++q) {
    GemmDesc d;
    d.transA = 0; d.transB = 1; d.M = n; d.N = n; d.K = n;
    d.A = Q + q * ss; d.lda = ls;
    d.B = Q + q * ss + n; d.ldb = ls;
    d.C = Z; d.ldc = ldz;
    d.alpha = 1.0; d.alphaVec = d_alpha + q; d.beta = q ? 1.0 : 0.0; d.aLo = 256; d.bLo = 256;
    ZOLO_CHECK(gemm(big_prec<T>(c), with_tile(c, d), c.s));
  }
  // Z = Q_X Z~
  ZOLO_CHECK(dense_apply_q<T>(c, Xw, n, 0, n, 1, Tx, Z, ldz, 0, ar));
  *nz = 1;
  return kOk;
}

#define ZOLO_LA_INST(T)                                                                                          \
  template int chol_inv<T>(const Ctx&, T*, int64_t, int64_t, T*, int64_t, int64_t, int64_t, T*, int64_t, int*, \
                           int);                                                                                 \
  template size_t chol_step_bytes<T>(int64_t, int);                                                              \
  template int chol_step<T>(const Ctx&, const T*, int64_t, int64_t, int, const double*, double, double,         \
                            const double*, T*, int64_t, int64_t, Arena&, int*, bool);                                \
  template size_t qr_step_bytes<T>(int64_t, int);                                                                \
  template int qr_step<T>(const Ctx&, const T*, int64_t, int64_t, int, const double*, const double*, T*, int64_t, \
                          int64_t, Arena&, int*, const QrPre<T>*);

ZOLO_LA_INST(double)
ZOLO_LA_INST(float)

}  // namespace zolo
