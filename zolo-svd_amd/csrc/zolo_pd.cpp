// Polar-decomposition drivers on one GPU: Zolo-PD (pzolopd1.f90:280-487) and
// QDWH-PD (pqdwhfacs.f90:296-466).
//
// The iteration logic (tolerances, QR-vs-Cholesky rule, coefficient and CON
// recurrences, FormX2 error) is the reference's; each iteration hands its r
// shifted factorisations to one batched set of launches (blockIdx.z = shift)
// and sums the shift contributions inside the fused FormX2 kernel, which
// replaces the DGSUM2D reduction across shift groups.  The only host round
// trip per iteration is the ERR scalar that the loop condition needs.
#include <atomic>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "coeff.hpp"
#include "driver.hpp"
#include "linalg.hpp"

namespace zolo {

namespace {
std::atomic<int> g_f64_products{1};
}  // namespace
int f64_products() { return g_f64_products.load(); }
void set_f64_products(int mode) { g_f64_products.store(mode); }


namespace {

constexpr int kPartials = 1024;
constexpr double kLowConverged = 6e-8;  // FP32 unit roundoff

struct Timer {
  hipStream_t s;
  std::vector<hipEvent_t> ev;
  std::vector<int> slot;
  explicit Timer(hipStream_t st) : s(st) {}
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  void mark(int phase) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    ev.push_back(e);
    slot.push_back(phase);
  }
  // phase_ms[slot[i]] += t(ev[i+1]) - t(ev[i]); total = last - first.
  void finish(zolo_info_t* info) {
    if (!info || ev.size() < 2) return;
    (void)hipEventSynchronize(ev.back());
    for (size_t i = 0; i + 1 < ev.size(); ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess && slot[i] >= 0 && slot[i] < 8)
        info->phase_ms[slot[i]] += ms;
    }
    float tot = 0.f;
    if (hipEventElapsedTime(&tot, ev.front(), ev.back()) == hipSuccess) info->time_ms = tot;
  }
};

enum Phase { kPhBounds = 0, kPhQr = 1, kPhChol = 2, kPhForm = 3, kPhPolish = 4, kPhH = 5, kPhOther = 6 };

// Small device-side scalars carved at the front of the workspace.
struct Scalars {
  double* shift;    // [8]
  double* alpha;    // [8]
  double* partial;  // [kPartials]
  double* out;      // [8]
  int* info;        // [8]
};

Scalars take_scalars(Arena& ar) {
  Scalars sc;
  sc.shift = ar.take<double>(8);
  sc.alpha = ar.take<double>(8);
  sc.partial = ar.take<double>(kPartials);
  sc.out = ar.take<double>(8);
  sc.info = ar.take<int>(8);
  return sc;
}
constexpr size_t kScalarBytes = 8 * 256 + (8 + 8 + kPartials + 8) * sizeof(double) + 8 * sizeof(int);

int fetch_double(const double* d, double* h, hipStream_t s) {
  ZOLO_HIP(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  return kOk;
}

}  // namespace

// H = (U^T A + A^T U)/2  (pqdwhfacs.f90:460-466, ZoloCHK1.f90:80-83)
int form_h(const Ctx& c, const double* U, int64_t ldu, const double* A, int64_t lda, int64_t n, double* H,
           int64_t ldh) {
  GemmDesc d;
  d.transA = 1; d.transB = 0; d.M = n; d.N = n; d.K = n;
  d.A = U; d.lda = ldu; d.B = A; d.ldb = lda; d.C = H; d.ldc = ldh;
  ZOLO_CHECK(gemm(c.big, d, c.s));
  return k_symmetrize<double>(H, ldh, n, c.s);
}

// ---------------------------------------------------------------------------
// Zolo-PD
// ---------------------------------------------------------------------------
int zolo_pd_run(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A, int64_t lda,
                double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info, hipStream_t s, int max_it,
                bool restart) {
  // alpha = beta = 0: the bounds come from the QR of A/alpha that the first
  // QR step needs anyway (see below); rk = 0 then picks choosem(1/beta)
  const bool autob = alpha == 0.0 && beta == 0.0;
  if (precision < 0 || precision > 4) return -1;
  // 3 and 4: FP32 / BF16x3 Cholesky steps on the reference's schedule (every
  // step takes all r shifts and the loop ends on the ERR test only)
  const bool refsched = precision >= 3;
  if (refsched) precision -= 2;
  if (sym != 'S' && sym != 's' && sym != 'N' && sym != 'n') return -2;
  if (rk < (autob ? 0 : 1) || rk > 8) return -3;
  if (n < 1) return -4;
  if (!autob && !(alpha > 0.0)) return -5;
  if (!autob && (!(beta > 0.0) || !(beta <= 1.0))) return -6;
  if (autob && restart) return -6;
  if (lda < n) return -8;
  if (ldu < n) return -10;
  if (H && ldh < n) return -12;
  const bool mixed = precision != 0;
  const bool symS = (sym == 'S' || sym == 's');
  zolo_info_t local;
  if (!info) info = &local;
  std::memset(info, 0, sizeof(*info));
  const int rcap = rk > 0 ? rk : 8;  // shifts the workspace is sized for

  const int64_t nn = n * n;
  const int64_t nout = (n + 255) / 256;
  // Workspace: scalars | X64 | Z64 (r) | [mixed: X32, Z32 (r), Xref64, polish] |
  // [auto bounds: the QR of A/alpha] | step scratch
  size_t step64 = std::max(qr_step_bytes<double>(n, rcap), chol_step_bytes<double>(n, rcap));
  if (restart) step64 = std::max(step64, bounds_bytes(n));
  if (autob) step64 = std::max(step64, dense_qr_bytes<double>(n) + rbound_bytes(n) + 4096);
  size_t bytes = kScalarBytes + sizeof(double) * (size_t)(1 + rcap) * nn + step64 + 4096;
  if (mixed) bytes += sizeof(float) * (size_t)(1 + rcap) * nn + sizeof(double) * nn + polish_bytes(n) +
                      chol_step_bytes<float>(n, rcap);
  if (autob) bytes += sizeof(double) * ((size_t)nn + (size_t)n + (size_t)nout * 256 * 256) + 3 * 256;
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  Scalars sc = take_scalars(ar);
  double* X = ar.take<double>(nn);
  double* Z = ar.take<double>((size_t)rcap * nn);
  float* X32 = nullptr;
  float* Z32 = nullptr;
  double* Xref = nullptr;
  char* polish_ws = nullptr;
  if (mixed) {
    X32 = ar.take<float>(nn);
    Z32 = ar.take<float>((size_t)rcap * nn);
    Xref = ar.take<double>(nn);
    polish_ws = ar.take<char>(polish_bytes(n));
  }
  QrPre<double> pre{nullptr, nullptr};
  double* pre_tau = nullptr;
  if (autob) {
    pre.Xw = ar.take<double>(nn);
    pre_tau = ar.take<double>(n);
    pre.Tx = ar.take<double>((size_t)nout * 256 * 256);
  }
  const size_t mark = ar.used();
  ZOLO_ARENA_OK(ar);

  Ctx c;
  c.s = s;
  c.low = (precision == 2) ? Prec::BF16X3 : Prec::F32;
  c.big = (mixed && f64_products() == 1) ? Prec::F64X : Prec::F64;
  Timer tm(s);
  tm.mark(kPhOther);

  if (autob) {
    // alpha = ||A||_F; A/alpha = Q_X R; beta = 1/(||R^{-1}||_1 est * sqrt(n)).
    // This is test_zolodrv1.f90:179-189's bound (rcond/sqrt(n) with ANORM =
    // 1) with the 1-norm estimate taken on R instead of on the LU factors:
    // sigma_min(A/alpha) = 1/||R^{-1}||_2 >= 1/(sqrt(n) ||R^{-1}||_1), so it
    // is a lower bound by the same argument.  The factorisation is the one the
    // first (QR) step of the iteration starts from, scaled by 1/beta.
    tm.mark(kPhBounds);
    ZOLO_CHECK(k_sumsq<double>(A, lda, n, n, sc.partial, kPartials, sc.out, s));
    double a2 = 0.0;
    ZOLO_CHECK(fetch_double(sc.out, &a2, s));
    alpha = std::sqrt(a2);
    if (!(alpha > 0.0) || !std::isfinite(alpha)) {
      set_error("zolo_pd: zero or non-finite matrix");
      return kErrArg;
    }
    ZOLO_CHECK(k_scale_copy<double, double>(A, lda, pre.Xw, n, n, n, 1.0 / alpha, s));
    tm.mark(kPhQr);
    ZOLO_CHECK(dense_geqrf<double>(c, pre.Xw, n, 0, n, 1, pre_tau, pre.Tx, ar));
    ar.reset();
    ar.take<char>(mark);
    tm.mark(kPhBounds);
    double est = 0.0;
    ZOLO_CHECK(rbound_run(n, pre.Xw, n, &est, ar, s));
    ar.reset();
    ar.take<char>(mark);
    if (!(est > 0.0) || !std::isfinite(est)) {
      set_error("zolo_pd: matrix is numerically singular (no positive lower bound from its R factor)");
      return kErrArg;
    }
    beta = std::min(1.0, 1.0 / (est * std::sqrt((double)n)));
    if (!(beta > 0.0)) {
      set_error("zolo_pd: lower bound underflows");
      return kErrArg;
    }
    if (rk == 0) rk = choosem(1.0 / beta);
    ZOLO_CHECK(k_scale_upper(pre.Xw, n, n, 1.0 / beta, s));
    tm.mark(kPhOther);
  }
  info->r = rk;
  info->alpha = alpha;
  info->beta = beta;

  ZOLO_HIP(hipMemsetAsync(sc.info, 0, 8 * sizeof(int), s));
  // X0 = A / (alpha*beta), CON = 1/beta  (pzolopd1.f90:286-291)
  ZOLO_CHECK(k_scale_copy<double, double>(A, lda, X, n, n, n, 1.0 / (alpha * beta), s));
  double con = 1.0 / beta;
  const double tol = dlamch_p();
  const double tol1 = 10.0 * dlamch_p() / 2.0;
  const double tol2 = std::exp(std::log(tol1) / (2.0 * rk + 1.0));
  int it = 0;
  double err = 1.0;
  bool low_active = false;   // iterate lives in X32
  bool have_ref = false;     // Xref holds the last FP64 iterate
  std::vector<double> coeff(3 * rk);
  double hs[8], ha[8];
  while (err > tol2 && it < max_it) {
    compute_coeff(con, tol, rk, coeff.data());
    double maxc = 0.0;
    for (int i = 0; i < 2 * rk; ++i) maxc = std::max(maxc, coeff[i]);
    if (it < 16) {
      info->con[it] = con;
      info->maxc[it] = maxc;
    }
    const bool use_qr = (it == 0 || con > 1e3) && maxc > 3e2;
    ar.reset();
    ar.take<char>(mark);
    // Shifts of this step.  The FP64 path keeps the reference's r.  A
    // low-precision Cholesky step takes the fewest shifts whose Zolotarev
    // bound reaches FP32 convergence in this one step (the polish then works
    // to FP64 from that iterate), falling back to r.
    int rs = rk;
    if (mixed && !refsched && !use_qr) {
      std::vector<double> cr(3 * rk);
      for (int r = 1; r < rk; ++r) {
        double f1 = 0.0, c1 = 0.0;
        compute_coeff(con, tol, r, cr.data());
        zolo_fone_con(con, cr.data(), r, &f1, &c1);
        if (c1 - 1.0 <= kLowConverged) {
          rs = r;
          std::copy(cr.begin(), cr.begin() + 3 * r, coeff.begin());
          break;
        }
      }
    }
    if (it < 16) info->shifts[it] = rs;
    int nz = rs;  // Z slots FormX2 sums (the factored QR step returns one)
    if (use_qr) {
      if (low_active) {  // cannot happen with the reference's rule after a Cholesky step; keep FP64 anyway
        ZOLO_CHECK(k_scale_copy<float, double>(X32, n, X, n, n, n, 1.0, s));
        low_active = false;
      }
      for (int q = 0; q < rk; ++q) {
        hs[q] = std::sqrt(coeff[2 * q]);
        ha[q] = -coeff[2 * rk + q] / hs[q];
      }
      ZOLO_HIP(hipMemcpyAsync(sc.shift, hs, rk * sizeof(double), hipMemcpyHostToDevice, s));
      ZOLO_HIP(hipMemcpyAsync(sc.alpha, ha, rk * sizeof(double), hipMemcpyHostToDevice, s));
      tm.mark(kPhQr);
      // the first step may start from the factorisation of the bounds
      const QrPre<double>* pq = (autob && it == 0) ? &pre : nullptr;
      ZOLO_CHECK(qr_step<double>(c, X, n, n, rk, sc.shift, sc.alpha, Z, n, nn, ar, &nz, pq));
      if (it < 16) info->schedule[it] = 'Q';
    } else {
      for (int q = 0; q < rs; ++q) {
        hs[q] = coeff[2 * q];
        ha[q] = -coeff[2 * rs + q];
      }
      ZOLO_HIP(hipMemcpyAsync(sc.shift, hs, rs * sizeof(double), hipMemcpyHostToDevice, s));
      ZOLO_HIP(hipMemcpyAsync(sc.alpha, ha, rs * sizeof(double), hipMemcpyHostToDevice, s));
      tm.mark(kPhChol);
      if (!mixed) {
        ZOLO_CHECK(chol_step<double>(c, X, n, n, rk, sc.shift, 1.0, 0.0, sc.alpha, Z, n, nn, ar, sc.info, symS));
        if (it < 16) info->schedule[it] = 'C';
      } else {
        if (!low_active) {
          // Keep the last FP64 iterate: it carries A's polar factor to FP64
          // accuracy and is what the final polish is measured against.
          ZOLO_HIP(hipMemcpyAsync(Xref, X, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
          have_ref = true;
          ZOLO_CHECK(k_scale_copy<double, float>(X, n, X32, n, n, n, 1.0, s));
          low_active = true;
        }
        ZOLO_CHECK(chol_step<float>(c, X32, n, n, rs, sc.shift, 1.0, 0.0, sc.alpha, Z32, n, nn, ar, sc.info, symS));
        if (it < 16) info->schedule[it] = 'c';
      }
    }
    double fone;
    zolo_fone_con(con, coeff.data(), rs, &fone, &con);
    tm.mark(kPhForm);
    if (low_active)
      ZOLO_CHECK(k_formx2<float>(X32, n, Z32, n, nn, nz, n, fone, sc.partial, kPartials, sc.out, s));
    else
      ZOLO_CHECK(k_formx2<double>(X, n, Z, n, nn, nz, n, fone, sc.partial, kPartials, sc.out, s));
    double e2;
    ZOLO_CHECK(fetch_double(sc.out, &e2, s));
    err = std::sqrt(e2);
    if (it < 16) info->err[it] = err;
    ++it;
    // PZOLOPD2's restart (pzolopd2.f90:397-405, 480-486): CON has reached 1
    // while ERR is still above TOL2, i.e. the scalar bound no longer describes
    // the iterate (first steps at CON near 1/u).  The reference hands this to
    // FormXTwo, which it never defines; here the bounds of the current iterate
    // are re-estimated (alpha, beta as test_zolodrv1.f90:179-189) and the
    // iteration continues from X / (alpha beta) with CON = 1/beta.
    if (restart && !low_active && con == 1.0 && err > tol2 && it < max_it) {
      ar.reset();
      ar.take<char>(mark);
      double ax = 0.0, bx = 0.0;
      ZOLO_CHECK(bounds_run(0, n, X, n, &ax, &bx, ar, s));
      if (bx > 0.0 && bx < 1.0) {
        ZOLO_CHECK(k_scale_copy<double, double>(X, n, X, n, n, n, 1.0 / (ax * bx), s));
        con = 1.0 / bx;
        ++info->restarts;
      }
    }
    // Mixed precision: once the Zolotarev bound says the low-precision iterate
    // has converged to its working precision (CON - 1 below the FP32 unit
    // roundoff), a further low-precision step cannot improve it; the FP64
    // polish takes over.  The FP64 path keeps the reference's ERR test.
    if (low_active && !refsched && con - 1.0 <= kLowConverged) break;
  }
  info->iters = it;
  int hinfo[8];
  ZOLO_HIP(hipMemcpyAsync(hinfo, sc.info, sizeof(hinfo), hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < rk; ++q)
    if (hinfo[q] && !info->info) info->info = hinfo[q];

  if (low_active) {
    tm.mark(kPhPolish);
    ZOLO_CHECK(k_scale_copy<float, double>(X32, n, X, n, n, n, 1.0, s));
    if (!have_ref) ZOLO_CHECK(k_scale_copy<double, double>(A, lda, Xref, n, n, n, 1.0, s));
    int passes = 0;
    ZOLO_CHECK(polish(c, X, n, Xref, n, n, polish_ws, &passes, info->cg_iters));
    info->polish_passes = passes;
  }
  tm.mark(kPhOther);
  ZOLO_CHECK(k_scale_copy<double, double>(X, n, U, ldu, n, n, 1.0, s));
  if (H) {
    tm.mark(kPhH);
    ZOLO_CHECK(form_h(c, U, ldu, A, lda, n, H, ldh));
  }
  tm.mark(kPhOther);
  tm.finish(info);
  if (info->info) {
    set_error("zolo_pd: shifted Gram matrix not positive definite");
    return kErrNotSpd;
  }
  return kOk;
}

// ---------------------------------------------------------------------------
// QDWH-PD (pqdwhfacs.f90:296-466) with alpha, beta from the GPU bounds kernel.
// ---------------------------------------------------------------------------
int qdwh_run(char sym, int64_t n, double sigma, const double* A, int64_t lda, double* U, int64_t ldu, double* H,
             int64_t ldh, zolo_info_t* info, hipStream_t s) {
  if (sym != 'S' && sym != 's' && sym != 'N' && sym != 'n') return -1;
  if (n < 1) return -2;
  if (lda < n) return -5;
  if (ldu < n) return -7;
  if (H && ldh < n) return -9;
  const bool symS = (sym == 'S' || sym == 's');
  zolo_info_t local;
  if (!info) info = &local;
  std::memset(info, 0, sizeof(*info));
  info->r = 1;
  const int64_t nn = n * n;
  size_t step = std::max(qr_step_bytes<double>(n, 1), chol_step_bytes<double>(n, 1));
  step = std::max(step, bounds_bytes(n));
  size_t bytes = kScalarBytes + sizeof(double) * 3 * nn + step + 4096;
  char* base = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &base));
  Arena ar(base, bytes);
  Scalars sc = take_scalars(ar);
  double* Q = ar.take<double>(nn);
  double* X1 = ar.take<double>(nn);
  double* Z = ar.take<double>(nn);
  const size_t mark = ar.used();
  Ctx c;
  c.s = s;
  Timer tm(s);
  tm.mark(kPhBounds);
  ZOLO_HIP(hipMemsetAsync(sc.info, 0, 8 * sizeof(int), s));
  // Q = A - sigma I ; alpha = ||Q||_F ; beta = ||Q/alpha||_1 * rcond / sqrt(n)
  ZOLO_CHECK(k_scale_copy<double, double>(A, lda, Q, n, n, n, 1.0, s));
  if (sigma != 0.0) ZOLO_CHECK(k_add_diag<double>(Q, n, n, -sigma, s));
  double alpha, beta;
  ZOLO_CHECK(bounds_run(1, n, Q, n, &alpha, &beta, ar, s));
  info->alpha = alpha;
  info->beta = beta;
  ZOLO_CHECK(k_scale_copy<double, double>(Q, n, Q, n, n, n, 1.0 / alpha, s));

  const double tol1 = 10.0 * dlamch_p() / 2.0;
  const double tol2 = std::exp(std::log(tol1) / 3.0);
  int it = 0;
  double lowk = beta, err = 1.0, err2 = std::fabs(1.0 - lowk);
  while ((it == 0 || err > tol2 || err2 > tol1) && it <= 12) {
    const double ak = hfunc(lowk);
    const double bk = (ak - 1.0) * (ak - 1.0) / 4.0;
    const double ck = ak + bk - 1.0;
    const double sck = std::sqrt(ck);
    const double lowk2 = lowk * lowk;
    if (it < 16) {
      info->con[it] = lowk;
      info->maxc[it] = ck;
    }
    lowk = lowk * (ak + bk * lowk2) / (1.0 + ck * lowk2);
    ar.reset();
    ar.take<char>(mark);
    double hs, ha;
    if (ck > 100) {
      hs = 1.0 / sck;
      ha = (ak - bk / ck) / sck;
      ZOLO_HIP(hipMemcpyAsync(sc.shift, &hs, sizeof(double), hipMemcpyHostToDevice, s));
      ZOLO_HIP(hipMemcpyAsync(sc.alpha, &ha, sizeof(double), hipMemcpyHostToDevice, s));
      tm.mark(kPhQr);
      int nz1 = 1;
      ZOLO_CHECK(qr_step<double>(c, Q, n, n, 1, sc.shift, sc.alpha, Z, n, nn, ar, &nz1));
      if (it < 16) info->schedule[it] = 'Q';
      if (it < 16) info->shifts[it] = 1;
    } else {
      hs = 0.0;
      ha = ak - bk / ck;
      ZOLO_HIP(hipMemcpyAsync(sc.shift, &hs, sizeof(double), hipMemcpyHostToDevice, s));
      ZOLO_HIP(hipMemcpyAsync(sc.alpha, &ha, sizeof(double), hipMemcpyHostToDevice, s));
      tm.mark(kPhChol);
      ZOLO_CHECK(chol_step<double>(c, Q, n, n, 1, sc.shift, ck, 1.0, sc.alpha, Z, n, nn, ar, sc.info, symS));
      if (it < 16) info->schedule[it] = 'C';
      if (it < 16) info->shifts[it] = 1;
    }
    // X1 = (b/c) Q + Z ; err = ||Q - X1||_F ; Q = X1
    tm.mark(kPhForm);
    ZOLO_CHECK(k_axpby<double>(bk / ck, Q, n, 1.0, Z, n, X1, n, n, n, s));
    ZOLO_CHECK(k_axpby<double>(1.0, Q, n, -1.0, X1, n, Z, n, n, n, s));
    ZOLO_CHECK(k_sumsq<double>(Z, n, n, n, sc.partial, kPartials, sc.out, s));
    std::swap(Q, X1);
    double e2;
    ZOLO_CHECK(fetch_double(sc.out, &e2, s));
    err = std::sqrt(e2);
    if (it < 16) info->err[it] = err;
    ++it;
    err2 = std::fabs(1.0 - lowk);
  }
  info->iters = it;
  int hinfo[8];
  ZOLO_HIP(hipMemcpyAsync(hinfo, sc.info, sizeof(hinfo), hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  info->info = hinfo[0];
  tm.mark(kPhOther);
  ZOLO_CHECK(k_scale_copy<double, double>(Q, n, U, ldu, n, n, 1.0, s));
  if (H) {
    tm.mark(kPhH);
    ZOLO_CHECK(form_h(c, U, ldu, A, lda, n, H, ldh));
  }
  tm.mark(kPhOther);
  tm.finish(info);
  if (info->info) {
    set_error("qdwh: shifted Gram matrix not positive definite");
    return kErrNotSpd;
  }
  return kOk;
}

}  // namespace zolo
