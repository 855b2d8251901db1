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
int gemm_t(const Ctx& c, GemmDesc d) {
  return gemm(prec_for<T>(c), d, c.s);
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
static int qr_nwg(int64_t mrows) {
  int64_t w = (mrows + 255) / 256;
  if (w > 256) w = 256;
  return (int)std::max<int64_t>(w, 1);
}

template <typename T>
static size_t stacked_qr_bytes(int64_t n, int r) {
  const int nb = kQrNb;
  const int64_t mr = n + nb;
  const int nwg = qr_nwg(mr);
  const int64_t npan = (n + nb - 1) / nb;
  size_t b = 0;
  b += sizeof(T) * (size_t)r * mr * nb;            // V
  b += sizeof(T) * (size_t)r * nb * nb;            // G
  b += sizeof(T) * (size_t)r * npan * nb * nb;     // T per panel
  b += 2 * sizeof(T) * (size_t)r * nb * n;         // W, W2
  b += sizeof(double) * (size_t)r * nb * ((size_t)nwg * nb + nb);  // panel partials
  b += sizeof(T) * (size_t)r * n;                  // tau
  return b + 8 * 256;
}

template <typename T>
int stacked_geqrf(const Ctx& c, T* S, int64_t lds, int64_t sS, int64_t n, int batch, T* tau, Arena& ar) {
  const int nb = kQrNb;
  const int64_t mr = n + nb;
  const int nwg_max = qr_nwg(mr);
  const int64_t npan = (n + nb - 1) / nb;
  T* V = ar.take<T>((size_t)batch * mr * nb);
  T* G = ar.take<T>((size_t)batch * nb * nb);
  T* Tp = ar.take<T>((size_t)batch * nb * nb);
  T* W = ar.take<T>((size_t)batch * nb * n);
  T* W2 = ar.take<T>((size_t)batch * nb * n);
  double* part = ar.take<double>((size_t)batch * nb * ((size_t)nwg_max * nb + nb));
  ZOLO_ARENA_OK(ar);
  const int64_t sV = mr * nb, sG = (int64_t)nb * nb, sW = (int64_t)nb * n;
  for (int64_t pi = 0; pi < npan; ++pi) {
    const int64_t p = pi * nb;
    const int jb = (int)std::min<int64_t>(nb, n - p);
    const int64_t mrows = n + jb;
    const int nwg = qr_nwg(mrows);
    for (int64_t j = p; j <= p + jb; ++j)
      ZOLO_CHECK(k_qr_panel_step<T>(S, lds, sS, n, p, jb, j, j > p, j < p + jb, part, nwg, tau, n, batch, c.s));
    const int64_t ncols = n - p - jb;
    if (ncols <= 0) continue;
    ZOLO_CHECK(k_qr_extract_v<T>(S, lds, sS, V, mrows, sV, p, jb, mrows, batch, c.s));
    GemmDesc d; d.batch = batch;           // G = V^T V
    d.transA = 1; d.transB = 0; d.M = jb; d.N = jb; d.K = mrows;
    d.A = V; d.lda = mrows; d.strideA = sV; d.B = V; d.ldb = mrows; d.strideB = sV;
    d.C = G; d.ldc = jb; d.strideC = sG;
    ZOLO_CHECK(gemm_t<T>(c, d));
    ZOLO_CHECK(k_qr_form_t<T>(G, jb, sG, tau, n, p, jb, Tp, jb, sG, batch, c.s));
    T* Cm = S + p + (p + jb) * lds;
    d = GemmDesc(); d.batch = batch;       // W = V^T C
    d.transA = 1; d.transB = 0; d.M = jb; d.N = ncols; d.K = mrows;
    d.A = V; d.lda = mrows; d.strideA = sV; d.B = Cm; d.ldb = lds; d.strideB = sS;
    d.C = W; d.ldc = jb; d.strideC = sW; d.aLo = 0;
    ZOLO_CHECK(gemm_t<T>(c, d));
    d = GemmDesc(); d.batch = batch;       // W2 = T^T W
    d.transA = 1; d.transB = 0; d.M = jb; d.N = ncols; d.K = jb;
    d.A = Tp; d.lda = jb; d.strideA = sG; d.B = W; d.ldb = jb; d.strideB = sW;
    d.C = W2; d.ldc = jb; d.strideC = sW;
    ZOLO_CHECK(gemm_t<T>(c, d));
    d = GemmDesc(); d.batch = batch;       // C -= V W2
    d.transA = 0; d.transB = 0; d.M = mrows; d.N = ncols; d.K = jb;
    d.A = V; d.lda = mrows; d.strideA = sV; d.B = W2; d.ldb = jb; d.strideB = sW;
    d.C = Cm; d.ldc = lds; d.strideC = sS; d.alpha = -1.0; d.beta = 1.0;
    ZOLO_CHECK(gemm_t<T>(c, d));
  }
  return kOk;
}

template <typename T>
int stacked_orgqr(const Ctx& c, const T* S, int64_t lds, int64_t sS, const T* tau, T* Q, int64_t ldq, int64_t sQ,
                  int64_t n, int batch, Arena& ar) {
  const int nb = kQrNb;
  const int64_t mr = n + nb;
  const int64_t npan = (n + nb - 1) / nb;
  T* V = ar.take<T>((size_t)batch * mr * nb);
  T* G = ar.take<T>((size_t)batch * nb * nb);
  T* Tp = ar.take<T>((size_t)batch * nb * nb);
  T* W = ar.take<T>((size_t)batch * nb * n);
  T* W2 = ar.take<T>((size_t)batch * nb * n);
  ZOLO_ARENA_OK(ar);
  const int64_t sV = mr * nb, sG = (int64_t)nb * nb, sW = (int64_t)nb * n;
  // Explicit Q = H_0 ... H_{n-1} [I; 0], panels applied last to first; the
  // block reflector T of each panel is re-formed from V^T V and tau.
  ZOLO_CHECK(k_set_identity_top<T>(Q, ldq, sQ, 2 * n, n, batch, c.s));
  for (int64_t pi = npan - 1; pi >= 0; --pi) {
    const int64_t p = pi * nb;
    const int jb = (int)std::min<int64_t>(nb, n - p);
    const int64_t mrows = n + jb;
    ZOLO_CHECK(k_qr_extract_v<T>(S, lds, sS, V, mrows, sV, p, jb, mrows, batch, c.s));
    GemmDesc d; d.batch = batch;           // G = V^T V
    d.transA = 1; d.transB = 0; d.M = jb; d.N = jb; d.K = mrows;
    d.A = V; d.lda = mrows; d.strideA = sV; d.B = V; d.ldb = mrows; d.strideB = sV;
    d.C = G; d.ldc = jb; d.strideC = sG;
    ZOLO_CHECK(gemm_t<T>(c, d));
    ZOLO_CHECK(k_qr_form_t<T>(G, jb, sG, tau, n, p, jb, Tp, jb, sG, batch, c.s));
    T* Qs = Q + p + p * ldq;
    const int64_t nc = n - p;
    d = GemmDesc(); d.batch = batch;       // W = V^T Qs
    d.transA = 1; d.transB = 0; d.M = jb; d.N = nc; d.K = mrows;
    d.A = V; d.lda = mrows; d.strideA = sV; d.B = Qs; d.ldb = ldq; d.strideB = sQ;
    d.C = W; d.ldc = jb; d.strideC = sW; d.aLo = 0;
    ZOLO_CHECK(gemm_t<T>(c, d));
    d = GemmDesc(); d.batch = batch;       // W2 = T W
    d.transA = 0; d.transB = 0; d.M = jb; d.N = nc; d.K = jb;
    d.A = Tp; d.lda = jb; d.strideA = sG; d.B = W; d.ldb = jb; d.strideB = sW;
    d.C = W2; d.ldc = jb; d.strideC = sW;
    ZOLO_CHECK(gemm_t<T>(c, d));
    d = GemmDesc(); d.batch = batch;       // Qs -= V W2
    d.transA = 0; d.transB = 0; d.M = mrows; d.N = nc; d.K = jb;
    d.A = V; d.lda = mrows; d.strideA = sV; d.B = W2; d.ldb = jb; d.strideB = sW;
    d.C = Qs; d.ldc = ldq; d.strideC = sQ; d.alpha = -1.0; d.beta = 1.0;
    ZOLO_CHECK(gemm_t<T>(c, d));
  }
  return kOk;
}

template <typename T>
int stacked_qr(const Ctx& c, T* S, int64_t lds, int64_t sS, T* Q, int64_t ldq, int64_t sQ, int64_t n, int batch,
               T* tau_out, Arena& ar) {
  T* tau = tau_out ? tau_out : ar.take<T>((size_t)batch * n);
  ZOLO_ARENA_OK(ar);
  const size_t mark = ar.used();
  ZOLO_CHECK(stacked_geqrf<T>(c, S, lds, sS, n, batch, tau, ar));
  if (!Q) return kOk;
  ar.reset();  // ORGQR reuses the GEQRF scratch
  ar.take<char>(mark);
  return stacked_orgqr<T>(c, S, lds, sS, tau, Q, ldq, sQ, n, batch, ar);
}

template <typename T>
size_t qr_step_bytes(int64_t n, int r) {
  return sizeof(T) * (size_t)4 * r * n * n + stacked_qr_bytes<T>(n, r) + 4096;
}

template <typename T>
int qr_step(const Ctx& c, const T* X, int64_t ldx, int64_t n, int r, const double* d_sdiag, const double* d_alpha,
            T* Z, int64_t ldz, int64_t sZ, Arena& ar) {
  const int64_t ls = 2 * n, ss = 2 * n * n;
  T* S = ar.take<T>((size_t)r * ss);
  T* Q = ar.take<T>((size_t)r * ss);
  ZOLO_ARENA_OK(ar);
  ZOLO_CHECK(k_build_stack<T>(X, ldx, S, ls, ss, n, d_sdiag, r, c.s));
  ZOLO_CHECK(stacked_qr<T>(c, S, ls, ss, Q, ls, ss, n, r, nullptr, ar));
  // Z_q = alpha_q * Q1 Q2^T ; Q2 is zero below its (kQrNb-1)-th subdiagonal.
  GemmDesc d; d.batch = r;
  d.transA = 0; d.transB = 1; d.M = n; d.N = n; d.K = n;
  d.A = Q; d.lda = ls; d.strideA = ss;
  d.B = Q + n; d.ldb = ls; d.strideB = ss;
  d.C = Z; d.ldc = ldz; d.strideC = sZ;
  d.alpha = 1.0; d.alphaVec = d_alpha; d.beta = 0.0; d.bLo = kQrNb;
  ZOLO_CHECK(gemm_t<T>(c, d));
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
                          int64_t, Arena&);                                                                      \
  template int stacked_qr<T>(const Ctx&, T*, int64_t, int64_t, T*, int64_t, int64_t, int64_t, int, T*, Arena&);   \
  template int stacked_geqrf<T>(const Ctx&, T*, int64_t, int64_t, int64_t, int, T*, Arena&);                      \
  template int stacked_orgqr<T>(const Ctx&, const T*, int64_t, int64_t, const T*, T*, int64_t, int64_t, int64_t, \
                                int, Arena&);
ZOLO_LA_INST(double)
ZOLO_LA_INST(float)

}  // namespace zolo
