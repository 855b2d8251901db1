// Two-level blocked structured Householder QR of the stacked matrix [X; s I]
// (mpdgeqrf.f:262-313, mpdorgqr.f:253-337), batched over the Zolotarev shifts.
//
// Inner level: panels of kQrNb = 64 columns are factored column by column
// (qr.hip; for small stacks each panel, 64 columns in FP32 or 32 in FP64, in
// a single workgroup) and update only the rest of their 256-wide outer block.  Outer
// level: the block reflector of all 256 columns, I - V T V^T, is assembled
// from the inner T factors (T_AB = -T_AA (V_A^T V_B) T_BB for adjacent column
// groups) and applied to the trailing matrix, and later to Q, as GEMMs with
// K = 256.  Every block reflector only touches rows p .. n+p+jb-1: rows below
// are still zero in the identity block of the stack.
#include <algorithm>

#include "eig.hpp"
#include "linalg.hpp"

namespace zolo {
namespace {

constexpr int kQrOuter = 256;

// Row ranges of the block reflector of columns [p, p + jb): [p, r1) of the
// top block and [n, r2) of the bottom block, stored contiguously in V.
struct Rows {
  int64_t r1, r2, len1, len2;
  int64_t m() const { return len1 + len2; }
};
inline Rows rows_of(QrShape sh, int64_t n, int64_t p, int64_t jb) {
  Rows r;
  r.r1 = (sh == QrShape::Tri) ? p + jb : n;
  r.r2 = (sh == QrShape::Dense) ? n : n + p + jb;
  r.len1 = r.r1 - p;
  r.len2 = r.r2 - n;
  return r;
}

template <typename T>
int gemm_t(const Ctx& c, const GemmDesc& d) {
  return gemm(prec_for<T>(c), with_tile(c, d), c.s);
}

// One thread per panel row (the panel kernel keeps each row in registers).
int qr_nwg(int64_t mrows) { return (int)std::max<int64_t>((mrows + 255) / 256, 1); }

// T of the outer block from its diagonal blk-blocks (already in Tb, ld
// kQrOuter) and G = V^T V: for adjacent column groups A | B,
// T_AB = -T_AA G_AB T_BB (the upper-triangular T of H_A H_B).
template <typename T>
int combine_t(const Ctx& c, T* Tb, int64_t sT, const T* G, int64_t sG, T* tmp, int64_t sTmp, int lo, int hi,
              int batch, int blk) {
  if (hi - lo <= blk) return kOk;
  int mid = lo + ((hi - lo) / 2 + blk - 1) / blk * blk;
  if (mid >= hi) mid = hi - blk;
  ZOLO_CHECK(combine_t<T>(c, Tb, sT, G, sG, tmp, sTmp, lo, mid, batch, blk));
  ZOLO_CHECK(combine_t<T>(c, Tb, sT, G, sG, tmp, sTmp, mid, hi, batch, blk));
  const int64_t ld = kQrOuter;
  const int a = mid - lo, b = hi - mid;
  GemmDesc d;  // tmp = G_AB T_BB   (T_BB upper: op(B)[k][j] = 0 for k > j)
  d.batch = batch;
  d.M = a; d.N = b; d.K = b;
  d.A = G + lo + (int64_t)mid * ld; d.lda = ld; d.strideA = sG;
  d.B = Tb + mid + (int64_t)mid * ld; d.ldb = ld; d.strideB = sT;
  d.C = tmp; d.ldc = a; d.strideC = sTmp; d.bHi = 0;
  ZOLO_CHECK(gemm_t<T>(c, d));
  d = GemmDesc();  // T_AB = -T_AA tmp   (T_AA upper: op(A)[i][k] = 0 for k < i)
  d.batch = batch;
  d.M = a; d.N = b; d.K = a;
  d.A = Tb + lo + (int64_t)lo * ld; d.lda = ld; d.strideA = sT;
  d.B = tmp; d.ldb = a; d.strideB = sTmp;
  d.C = Tb + lo + (int64_t)mid * ld; d.ldc = ld; d.strideC = sT; d.alpha = -1.0; d.aLo = 0;
  ZOLO_CHECK(gemm_t<T>(c, d));
  return kOk;
}

struct QrScratch {
  void* V;
  void* G;
  void* tmp;
  void* W;
  void* W2;
};

template <typename T>
int take_scratch(Arena& ar, int64_t n, int batch, QrScratch& q) {
  const int64_t nbo = kQrOuter, mr = n + nbo;
  q.V = ar.take<T>((size_t)batch * mr * nbo);
  q.G = ar.take<T>((size_t)batch * nbo * nbo);
  q.tmp = ar.take<T>((size_t)batch * nbo * nbo);
  q.W = ar.take<T>((size_t)batch * nbo * n);
  q.W2 = ar.take<T>((size_t)batch * nbo * n);
  ZOLO_ARENA_OK(ar);
  return kOk;
}

// Apply H^T = (I - V T V^T)^T (transpose) or H (left) to the ncols columns at
// Cm (row p of the matrix).  The reflector's rows are [0, len1) at Cm and
// [0, len2) at Cm + off2 (off2 = n - p); V holds them contiguously.  Its
// bottom part is zero for bottom row k > p + column (the fill-in structure).
template <typename T>
int apply_block_reflector(const Ctx& c, const T* V, int64_t ldv, int64_t sV, const T* Tm, int64_t ldt, int64_t sT,
                          int jb, T* Cm, int64_t ldc, int64_t sC, int64_t len1, int64_t off2, int64_t len2, int64_t p,
                          int64_t ncols, bool transpose, T* W, T* W2, int64_t sW, int batch) {
  const bool one = (off2 == len1) || len2 == 0;  // a single contiguous row range
  const int64_t m1 = (off2 == len1) ? len1 + len2 : len1;
  const Prec pr = prec_for<T>(c);
  // two row ranges in one launch where the engine takes them (FP64 / FP32):
  // rows [len1, len1 + len2) of V are rows off2 .. of C
  const bool gapB = !one && (pr == Prec::F64 || pr == Prec::F32) && len1 % 16 == 0;
  const bool gapC = gapB && len1 % 128 == 0;
  GemmDesc d;  // W = V^T C
  d.batch = batch;
  d.transA = 1; d.M = jb; d.N = ncols; d.K = gapB ? len1 + len2 : m1;
  d.A = V; d.lda = ldv; d.strideA = sV; d.B = Cm; d.ldb = ldc; d.strideB = sC;
  d.C = W; d.ldc = jb; d.strideC = sW; d.aLo = 0;
  if (gapB) {  // V2[k][i] = 0 for k > p + i: in the joined depth, k > len1 + p + i
    d.aLo = -1;
    d.aHi = (int)(len1 + p);
    d.bGapAt = len1;
    d.bGap = off2 - len1;
  }
  ZOLO_CHECK(gemm_t<T>(c, d));
  if (!one && !gapB) {  // W += V2^T C2   (V2[k][i] = 0 for k > p + i)
    d = GemmDesc();
    d.batch = batch;
    d.transA = 1; d.M = jb; d.N = ncols; d.K = len2;
    d.A = V + len1; d.lda = ldv; d.strideA = sV; d.B = Cm + off2; d.ldb = ldc; d.strideB = sC;
    d.C = W; d.ldc = jb; d.strideC = sW; d.beta = 1.0; d.aHi = (int)p;
    ZOLO_CHECK(gemm_t<T>(c, d));
  }
  d = GemmDesc();  // W2 = T^T W (GEQRF) or T W (ORGQR)
  d.batch = batch;
  d.transA = transpose ? 1 : 0; d.M = jb; d.N = ncols; d.K = jb;
  d.A = Tm; d.lda = ldt; d.strideA = sT; d.B = W; d.ldb = jb; d.strideB = sW;
  d.C = W2; d.ldc = jb; d.strideC = sW;
  // op(A) = T^T is lower triangular (zero for k > i), op(A) = T upper (zero for k < i)
  if (transpose) d.aHi = 0; else d.aLo = 0;
  ZOLO_CHECK(gemm_t<T>(c, d));
  d = GemmDesc();  // C -= V W2
  d.batch = batch;
  d.M = gapC ? len1 + len2 : m1; d.N = ncols; d.K = jb;
  d.A = V; d.lda = ldv; d.strideA = sV; d.B = W2; d.ldb = jb; d.strideB = sW;
  d.C = Cm; d.ldc = ldc; d.strideC = sC; d.alpha = -1.0; d.beta = 1.0;
  if (gapC) {  // V2[i][k] = 0 for i > p + k: in the joined rows, k < i - len1 - p
    d.aLo = (int)(len1 + p);
    d.cGapAt = len1;
    d.cGap = off2 - len1;
  }
  ZOLO_CHECK(gemm_t<T>(c, d));
  if (!one && !gapC) {  // C2 -= V2 W2
    d = GemmDesc();
    d.batch = batch;
    d.M = len2; d.N = ncols; d.K = jb;
    d.A = V + len1; d.lda = ldv; d.strideA = sV; d.B = W2; d.ldb = jb; d.strideB = sW;
    d.C = Cm + off2; d.ldc = ldc; d.strideC = sC; d.alpha = -1.0; d.beta = 1.0; d.aLo = (int)p;
    ZOLO_CHECK(gemm_t<T>(c, d));
  }
  return kOk;
}

template <typename T>
int geqrf_impl(const Ctx& c, T* S, int64_t lds, int64_t sS, int64_t n, int batch, T* tau, T* Tall, Arena& ar,
               QrShape sh, int64_t kc = -1) {
  // kc: columns to factor (the Dense shape may stop early: an n x kc matrix)
  if (kc < 0 || sh != QrShape::Dense) kc = n;
  // small stacks: every inner panel in one workgroup (narrower panels in FP64)
  const bool fused = sh == QrShape::Stack && n + kQrPanelWgCols<T> <= kQrPanelWgRows;
  const int nb = fused ? kQrPanelWgCols<T> : kQrNb, nbo = kQrOuter;
  QrScratch q;
  ZOLO_CHECK(take_scratch<T>(ar, n, batch, q));
  const int nwg_max = qr_nwg(n + nb);
  // slot records of every panel column (per shift), then 128 words per shift
  const int64_t part_doubles = (int64_t)batch * nb * ((int64_t)nwg_max * nb + nb) + (int64_t)batch * 64;
  double* part = ar.take<double>((size_t)part_doubles);
  ZOLO_ARENA_OK(ar);
  T* V = (T*)q.V;
  T* G = (T*)q.G;
  const int64_t mr = n + nbo;
  const int64_t sV = mr * nbo, sG = (int64_t)nbo * nbo, sW = (int64_t)nbo * n;
  const int64_t nout = (n + nbo - 1) / nbo;
  const int64_t sTall = nout * sG;
  for (int64_t P = 0; P < kc; P += nbo) {
    const int JB = (int)std::min<int64_t>(nbo, kc - P);
    T* Tb = Tall + (P / nbo) * sG;
    ZOLO_CHECK(k_fill<T>(Tb, nbo, sTall, nbo, nbo, 0.0, batch, c.s));
    for (int64_t p = P; p < P + JB; p += nb) {
      const int jb = (int)std::min<int64_t>(nb, P + JB - p);
      const Rows ri = rows_of(sh, n, p, jb);
      const int nwg = qr_nwg(ri.m());
      if (fused) {
        ZOLO_CHECK(k_qr_panel_wg<T>(S, lds, sS, n, p, jb, tau, n, batch, c.s));
      } else {
        bool done = false;
        ZOLO_CHECK(k_qr_panel_coop<T>(S, lds, sS, n, p, jb, ri.r1, ri.r2, part, part_doubles, tau, n, batch, c.s,
                                      &done));
        for (int64_t j = p; !done && j <= p + jb; ++j)
          ZOLO_CHECK(k_qr_panel_step<T>(S, lds, sS, n, p, jb, ri.r1, ri.r2, j, j > p, j < p + jb, part, nwg, tau, n,
                                        batch, c.s));
      }
      ZOLO_CHECK(k_qr_extract_v<T>(S, lds, sS, V, ri.m(), sV, p, jb, n, ri.r1, ri.r2, batch, c.s));
      GemmDesc d;  // G = V^T V -> T of the panel, stored on the diagonal of the outer T
      d.batch = batch;
      d.transA = 1; d.M = jb; d.N = jb; d.K = ri.m();
      d.A = V; d.lda = ri.m(); d.strideA = sV; d.B = V; d.ldb = ri.m(); d.strideB = sV;
      d.C = G; d.ldc = jb; d.strideC = sG;
      ZOLO_CHECK(gemm_t<T>(c, d));
      T* Tp = Tb + (p - P) * (nbo + 1);
      ZOLO_CHECK(k_qr_form_t<T>(G, jb, sG, tau, n, p, jb, Tp, nbo, sTall, batch, c.s));
      const int64_t ncols = P + JB - p - jb;  // rest of the outer block only
      if (ncols > 0)
        ZOLO_CHECK(apply_block_reflector<T>(c, V, ri.m(), sV, Tp, nbo, sTall, jb, S + p + (p + jb) * lds, lds, sS,
                                            ri.len1, n - p, ri.len2, p, ncols, true, (T*)q.W, (T*)q.W2, sW, batch));
    }
    const Rows ro = rows_of(sh, n, P, JB);
    ZOLO_CHECK(k_qr_extract_v<T>(S, lds, sS, V, ro.m(), sV, P, JB, n, ro.r1, ro.r2, batch, c.s));
    if (JB > nb) {
      GemmDesc d;
      d.batch = batch;
      d.transA = 1; d.M = JB; d.N = JB; d.K = ro.m();
      d.A = V; d.lda = ro.m(); d.strideA = sV; d.B = V; d.ldb = ro.m(); d.strideB = sV;
      d.C = G; d.ldc = nbo; d.strideC = sG;
      ZOLO_CHECK(gemm_t<T>(c, d));
      ZOLO_CHECK(combine_t<T>(c, Tb, sTall, G, sG, (T*)q.tmp, sG, 0, JB, batch, nb));
    }
    const int64_t ncols = kc - P - JB;
    if (ncols > 0)
      ZOLO_CHECK(apply_block_reflector<T>(c, V, ro.m(), sV, Tb, nbo, sTall, JB, S + P + (P + JB) * lds, lds, sS,
                                          ro.len1, n - P, ro.len2, P, ncols, true, (T*)q.W, (T*)q.W2, sW, batch));
  }
  return kOk;
}

template <typename T>
int orgqr_impl(const Ctx& c, const T* S, int64_t lds, int64_t sS, const T* tau, T* Q, int64_t ldq, int64_t sQ,
               int64_t n, int batch, T* Tall, bool have_t, Arena& ar, QrShape sh) {
  const int nb = kQrNb, nbo = kQrOuter;
  QrScratch q;
  ZOLO_CHECK(take_scratch<T>(ar, n, batch, q));
  T* V = (T*)q.V;
  T* G = (T*)q.G;
  const int64_t mr = n + nbo;
  const int64_t sV = mr * nbo, sG = (int64_t)nbo * nbo, sW = (int64_t)nbo * n;
  const int64_t nout = (n + nbo - 1) / nbo;
  const int64_t sTall = nout * sG;
  // Q = H_0 ... H_{n-1} [I; 0]: one outer block reflector at a time, last to first.
  ZOLO_CHECK(k_set_identity_top<T>(Q, ldq, sQ, 2 * n, n, batch, c.s));
  for (int64_t ob = nout - 1; ob >= 0; --ob) {
    const int64_t P = ob * nbo;
    const int JB = (int)std::min<int64_t>(nbo, n - P);
    const Rows ro = rows_of(sh, n, P, JB);
    T* Tb = Tall + ob * sG;
    ZOLO_CHECK(k_qr_extract_v<T>(S, lds, sS, V, ro.m(), sV, P, JB, n, ro.r1, ro.r2, batch, c.s));
    if (!have_t) {  // re-form T of the outer block from V and tau
      ZOLO_CHECK(k_fill<T>(Tb, nbo, sTall, nbo, nbo, 0.0, batch, c.s));
      GemmDesc d;
      d.batch = batch;
      d.transA = 1; d.M = JB; d.N = JB; d.K = ro.m();
      d.A = V; d.lda = ro.m(); d.strideA = sV; d.B = V; d.ldb = ro.m(); d.strideB = sV;
      d.C = G; d.ldc = nbo; d.strideC = sG;
      ZOLO_CHECK(gemm_t<T>(c, d));
      for (int64_t p = P; p < P + JB; p += nb) {
        const int jb = (int)std::min<int64_t>(nb, P + JB - p);
        const int64_t o = p - P;
        ZOLO_CHECK(k_qr_form_t<T>(G + o * (nbo + 1), nbo, sG, tau, n, p, jb, Tb + o * (nbo + 1), nbo, sTall, batch,
                                  c.s));
      }
      ZOLO_CHECK(combine_t<T>(c, Tb, sTall, G, sG, (T*)q.tmp, sG, 0, JB, batch, kQrNb));
    }
    ZOLO_CHECK(apply_block_reflector<T>(c, V, ro.m(), sV, Tb, nbo, sTall, JB, Q + P + P * ldq, ldq, sQ, ro.len1, n - P,
                                        ro.len2, P, n - P, false, (T*)q.W, (T*)q.W2, sW, batch));
  }
  return kOk;
}

// M <- Q M for the Q of a dense n x n GEQRF (reflectors and outer-block T
// factors from geqrf_impl), M n x n: the outer blocks from last to first.
template <typename T>
int apply_q_dense_impl(const Ctx& c, const T* S, int64_t lds, int64_t sS, const T* Tall, T* M, int64_t ldm,
                       int64_t sM, int64_t n, int batch, Arena& ar) {
  const int nbo = kQrOuter;
  QrScratch q;
  ZOLO_CHECK(take_scratch<T>(ar, n, batch, q));
  T* V = (T*)q.V;
  const int64_t mr = n + nbo;
  const int64_t sV = mr * nbo, sG = (int64_t)nbo * nbo, sW = (int64_t)nbo * n;
  const int64_t nout = (n + nbo - 1) / nbo;
  const int64_t sTall = nout * sG;
  for (int64_t ob = nout - 1; ob >= 0; --ob) {
    const int64_t P = ob * nbo;
    const int JB = (int)std::min<int64_t>(nbo, n - P);
    const Rows ro = rows_of(QrShape::Dense, n, P, JB);
    ZOLO_CHECK(k_qr_extract_v<T>(S, lds, sS, V, ro.m(), sV, P, JB, n, ro.r1, ro.r2, batch, c.s));
    ZOLO_CHECK(apply_block_reflector<T>(c, V, ro.m(), sV, Tall + ob * sG, nbo, sTall, JB, M + P, ldm, sM, ro.len1,
                                        n - P, 0, P, n, false, (T*)q.W, (T*)q.W2, sW, batch));
  }
  return kOk;
}

// Q (n x m) = the first m columns of H_0 ... H_{k-1} from a dense n x k GEQRF:
// Q = [I; 0], then the outer blocks from last to first, each on columns P ..
// m-1 only (columns before P are still unit vectors above its rows).
template <typename T>
int orgqr_dense_cols_impl(const Ctx& c, const T* A, int64_t lda, int64_t n, int64_t k, const T* Tall, T* Q,
                          int64_t ldq, int64_t m, Arena& ar) {
  const int nbo = kQrOuter;
  QrScratch q;
  ZOLO_CHECK(take_scratch<T>(ar, n, 1, q));
  T* V = (T*)q.V;
  const int64_t mr = n + nbo;
  const int64_t sV = mr * nbo, sG = (int64_t)nbo * nbo, sW = (int64_t)nbo * n;
  const int64_t nout = (k + nbo - 1) / nbo;
  ZOLO_CHECK(k_set_identity_top<T>(Q, ldq, 0, n, m, 1, c.s));
  for (int64_t ob = nout - 1; ob >= 0; --ob) {
    const int64_t P = ob * nbo;
    if (P >= m) continue;
    const int JB = (int)std::min<int64_t>(nbo, k - P);
    const Rows ro = rows_of(QrShape::Dense, n, P, JB);
    ZOLO_CHECK(k_qr_extract_v<T>(A, lda, 0, V, ro.m(), sV, P, JB, n, ro.r1, ro.r2, 1, c.s));
    ZOLO_CHECK(apply_block_reflector<T>(c, V, ro.m(), sV, Tall + ob * sG, nbo, nout * sG, JB, Q + P + P * ldq, ldq,
                                        0, ro.len1, n - P, 0, P, m - P, false, (T*)q.W, (T*)q.W2, sW, 1));
  }
  return kOk;
}

}  // namespace

template <typename T>
int dense_geqrf_cols(const Ctx& c, T* A, int64_t lda, int64_t n, int64_t k, T* tau, T* Tall, Arena& ar) {
  if (k < 1 || k > n) {
    set_error("dense QR: column count out of range");
    return kErrArg;
  }
  return geqrf_impl<T>(c, A, lda, 0, n, 1, tau, Tall, ar, QrShape::Dense, k);
}

template <typename T>
int dense_orgqr_cols(const Ctx& c, const T* A, int64_t lda, int64_t n, int64_t k, const T* Tall, T* Q, int64_t ldq,
                     int64_t m, Arena& ar) {
  if (k < 1 || k > n || m < 1 || m > n) {
    set_error("dense ORGQR: column counts out of range");
    return kErrArg;
  }
  return orgqr_dense_cols_impl<T>(c, A, lda, n, k, Tall, Q, ldq, m, ar);
}
template int dense_geqrf_cols<double>(const Ctx&, double*, int64_t, int64_t, int64_t, double*, double*, Arena&);
template int dense_orgqr_cols<double>(const Ctx&, const double*, int64_t, int64_t, int64_t, const double*, double*,
                                      int64_t, int64_t, Arena&);

int apply_trd_block(const Ctx& c, const double* A, int64_t lda, const double* tau, int64_t k0, int jb, int64_t n,
                    double* C, int64_t ldc, int64_t ncols, Arena& ar, int off) {
  const int nbo = kQrOuter;
  if (jb <= 0 || jb > nbo) {
    set_error("apply_trd_block: block width out of range");
    return kErrArg;
  }
  const int64_t r0 = k0 + off, mrows = n - r0;
  if (mrows <= 0) return kOk;
  double* V = ar.take<double>((size_t)mrows * jb);
  double* G = ar.take<double>((size_t)nbo * nbo);
  double* Tm = ar.take<double>((size_t)nbo * nbo);
  double* tmp = ar.take<double>((size_t)nbo * nbo);
  double* W = ar.take<double>((size_t)nbo * ncols);
  double* W2 = ar.take<double>((size_t)nbo * ncols);
  ZOLO_ARENA_OK(ar);
  ZOLO_CHECK(ke_trd_extract_v(A, lda, k0, r0, jb, mrows, V, mrows, off, c.s));
  GemmDesc d;  // G = V^T V
  d.transA = 1; d.M = jb; d.N = jb; d.K = mrows;
  d.A = V; d.lda = mrows; d.B = V; d.ldb = mrows; d.C = G; d.ldc = nbo;
  ZOLO_CHECK(gemm_t<double>(c, d));
  ZOLO_CHECK(k_fill<double>(Tm, nbo, 0, nbo, nbo, 0.0, 1, c.s));
  for (int o = 0; o < jb; o += kQrNb) {
    const int b = std::min(kQrNb, jb - o);
    ZOLO_CHECK(k_qr_form_t<double>(G + (int64_t)o * (nbo + 1), nbo, 0, tau, 0, k0 + o, b, Tm + (int64_t)o * (nbo + 1),
                                   nbo, 0, 1, c.s));
  }
  ZOLO_CHECK(combine_t<double>(c, Tm, 0, G, 0, tmp, 0, 0, jb, 1, kQrNb));
  return apply_block_reflector<double>(c, V, mrows, 0, Tm, nbo, 0, jb, C + r0, ldc, 0, mrows, mrows, 0, 0, ncols, false,
                                       W, W2, 0, 1);
}

template <typename T>
size_t stacked_qr_bytes(int64_t n, int r) {
  const int64_t nbo = kQrOuter, mr = n + nbo;
  const int nwg = qr_nwg(n + kQrNb);
  const int64_t nout = (n + nbo - 1) / nbo;
  size_t b = 0;
  b += sizeof(T) * (size_t)r * mr * nbo;                               // V
  b += 2 * sizeof(T) * (size_t)r * nbo * nbo;                          // G, combine scratch
  b += sizeof(T) * (size_t)r * nout * nbo * nbo;                       // T of every outer block
  b += 2 * sizeof(T) * (size_t)r * nbo * n;                            // W, W2
  b += sizeof(double) * (size_t)r * (kQrNb * ((size_t)nwg * kQrNb + kQrNb) + 64);  // panel partials, XCC words, flags
  b += sizeof(T) * (size_t)r * n;                                      // tau
  return b + 16 * 256;
}

template <typename T>
int stacked_geqrf(const Ctx& c, T* S, int64_t lds, int64_t sS, int64_t n, int batch, T* tau, Arena& ar) {
  const int64_t nout = (n + kQrOuter - 1) / kQrOuter;
  T* Tall = ar.take<T>((size_t)batch * nout * kQrOuter * kQrOuter);
  ZOLO_ARENA_OK(ar);
  return geqrf_impl<T>(c, S, lds, sS, n, batch, tau, Tall, ar, QrShape::Stack);
}

template <typename T>
int stacked_orgqr(const Ctx& c, const T* S, int64_t lds, int64_t sS, const T* tau, T* Q, int64_t ldq, int64_t sQ,
                  int64_t n, int batch, Arena& ar) {
  const int64_t nout = (n + kQrOuter - 1) / kQrOuter;
  T* Tall = ar.take<T>((size_t)batch * nout * kQrOuter * kQrOuter);
  ZOLO_ARENA_OK(ar);
  return orgqr_impl<T>(c, S, lds, sS, tau, Q, ldq, sQ, n, batch, Tall, false, ar, QrShape::Stack);
}

template <typename T>
int stacked_qr(const Ctx& c, T* S, int64_t lds, int64_t sS, T* Q, int64_t ldq, int64_t sQ, int64_t n, int batch,
               T* tau_out, Arena& ar, QrShape sh) {
  T* tau = tau_out ? tau_out : ar.take<T>((size_t)batch * n);
  const int64_t nout = (n + kQrOuter - 1) / kQrOuter;
  T* Tall = ar.take<T>((size_t)batch * nout * kQrOuter * kQrOuter);
  ZOLO_ARENA_OK(ar);
  const size_t mark = ar.used();
  ZOLO_CHECK(geqrf_impl<T>(c, S, lds, sS, n, batch, tau, Tall, ar, sh));
  if (!Q) return kOk;
  ar.reset();  // ORGQR reuses the GEQRF scratch; the outer-block T factors are kept
  ar.take<char>(mark);
  return orgqr_impl<T>(c, S, lds, sS, tau, Q, ldq, sQ, n, batch, Tall, true, ar, sh);
}

template <typename T>
size_t dense_qr_bytes(int64_t n, int batch) {
  const int64_t nout = (n + kQrOuter - 1) / kQrOuter;
  return stacked_qr_bytes<T>(n, batch) + sizeof(T) * (size_t)batch * ((size_t)nout * kQrOuter * kQrOuter + n) + 512;
}

template <typename T>
int dense_geqrf(const Ctx& c, T* A, int64_t lda, int64_t sA, int64_t n, int batch, T* tau, T* Tall, Arena& ar) {
  return geqrf_impl<T>(c, A, lda, sA, n, batch, tau, Tall, ar, QrShape::Dense);
}

template <typename T>
int dense_apply_q(const Ctx& c, const T* A, int64_t lda, int64_t sA, int64_t n, int batch, const T* Tall, T* M,
                  int64_t ldm, int64_t sM, Arena& ar) {
  return apply_q_dense_impl<T>(c, A, lda, sA, Tall, M, ldm, sM, n, batch, ar);
}

#define ZOLO_QRB_INST(T)                                                                                            \
  template size_t stacked_qr_bytes<T>(int64_t, int);                                                                \
  template int stacked_qr<T>(const Ctx&, T*, int64_t, int64_t, T*, int64_t, int64_t, int64_t, int, T*, Arena&,      \
                             QrShape);                                                                             \
  template size_t dense_qr_bytes<T>(int64_t, int);                                                                  \
  template int dense_geqrf<T>(const Ctx&, T*, int64_t, int64_t, int64_t, int, T*, T*, Arena&);                     \
  template int dense_apply_q<T>(const Ctx&, const T*, int64_t, int64_t, int64_t, int, const T*, T*, int64_t,       \
                                int64_t, Arena&);                                                                  \
  template int stacked_geqrf<T>(const Ctx&, T*, int64_t, int64_t, int64_t, int, T*, Arena&);                        \
  template int stacked_orgqr<T>(const Ctx&, const T*, int64_t, int64_t, const T*, T*, int64_t, int64_t, int64_t,   \
                                int, Arena&);
ZOLO_QRB_INST(double)
ZOLO_QRB_INST(float)

}  // namespace zolo
