// Symmetric eigensolver driver (the ELPA_PDSYEVD2 step of the Zolo-SVD tail,
// ZoloCHK1.f90:86) and the Zolo-SVD assembly (ZoloCHK1.f90:80-91): polar
// factor U_p, H = (U_p^T A + A^T U_p)/2, H = V diag(sigma) V^T, U = U_p V.
//
// The eigensolver: (1) blocked Householder tridiagonalisation (dsytrd with
// dlatrd panels, rank-2k trailing updates as two MFMA GEMMs on the full
// symmetric matrix); (2) Cuppen divide and conquer on the tridiagonal, level
// by level from 1 x 1 leaves, each merge's eigenvector update an MFMA GEMM;
// (3) back-transformation with the tridiagonalisation reflectors in blocks of
// 256 (compact WY).  Everything stays on the device; the host only plans the
// merge tree.
#include <chrono>
#include <vector>

#include "driver.hpp"
#include "eig.hpp"

namespace zolo {

namespace {

struct MergeH {
  int off, s1, s2;
};

struct Level {
  std::vector<MergeH> merges;
  int carry_off = -1, carry_s = 0;  // a block passed up unmerged
  size_t first = 0;                 // index of the level's first merge in the device array
};

std::vector<Level> plan_levels(int64_t n) {
  std::vector<Level> lv;
  std::vector<std::pair<int, int>> blocks;
  for (int64_t i = 0; i < n; ++i) blocks.push_back({(int)i, 1});
  size_t first = 0;
  while (blocks.size() > 1) {
    Level L;
    L.first = first;
    std::vector<std::pair<int, int>> next;
    size_t b = 0;
    for (; b + 1 < blocks.size(); b += 2) {
      L.merges.push_back({blocks[b].first, blocks[b].second, blocks[b + 1].second});
      next.push_back({blocks[b].first, blocks[b].second + blocks[b + 1].second});
    }
    if (b < blocks.size()) {
      L.carry_off = blocks[b].first;
      L.carry_s = blocks[b].second;
      next.push_back(blocks[b]);
    }
    first += L.merges.size();
    lv.push_back(std::move(L));
    blocks.swap(next);
  }
  return lv;
}

// ---------------------------------------------------------------------------
// Two-stage reduction (the ELPA 2-stage scheme of elpa_pdsyevd2.F90:199-200).
// ---------------------------------------------------------------------------
constexpr int64_t kTwoStageMin = 4 * kSbNb;  // smaller matrices take the one-stage reduction

struct TwoStage {
  int64_t n = 0, ldb = 0, ldzp = 0;
  int J = 0, kmax = 0, ngroups = 0;
  int64_t span = 0;
  int64_t ncols1 = 0;                   // columns whose stage-1 reflectors are stored in H
  double* tau1 = nullptr;               // stage-1 tau per column
  double *LB = nullptr, *vstore = nullptr, *tstore = nullptr, *Gs = nullptr, *Ts = nullptr;
  int *progress = nullptr, *status = nullptr;
  double *Zp = nullptr, *W = nullptr, *W2 = nullptr;
};

size_t two_stage_bytes(int64_t n) {
  const int64_t b = kSbNb, g = kSbGroup, span = b + g - 1;
  const int64_t J = (n - 2 + g - 1) / g, kmax = (n - 2) / b + 1, ng = J * kmax;
  size_t by = 0;
  by += sizeof(double) * (size_t)n;                                   // tau1
  by += sizeof(double) * (size_t)n * 2 * b;                           // LB
  by += sizeof(double) * (size_t)ng * g * (span + 1 + 2 * g);         // vstore, tstore, G, T
  by += sizeof(int) * ((size_t)n + 64);                               // progress, status
  by += sizeof(double) * (size_t)(n + span + b + g) * n;              // padded eigenvectors
  by += 2 * sizeof(double) * (size_t)J * g * n;                       // W, W2 of one wavefront
  // stage 1: panel QR scratch and the block vectors of the two-sided update
  by += dense_qr_bytes<double>(n) + sizeof(double) * (size_t)(4 * n * b + 4 * b * b + 256 * 256);
  return by + 64 * 256;
}

// Full -> band (bandwidth b): for every panel of b columns, Householder QR of
// the part below the band, then the two-sided update of the trailing matrix
// A22 <- Q^T A22 Q with Q = I - V T V^T:  Y = A22 V T, M = V^T Y,
// Y <- Y - V (T^T M)/2, A22 <- A22 - V Y^T - Y V^T (one product of K = 2b).
// The reflectors stay below R in H; their tau in ts.tau1.
int sy2sb(const Ctx& c, double* H, int64_t ldh, int64_t n, TwoStage& ts, Arena& ar) {
  const int b = kSbNb;
  const size_t mark = ar.used();
  ts.ncols1 = 0;
  for (int64_t k0 = 0; k0 + b + 1 < n; k0 += b) {
    ar.reset();
    ar.take<char>(mark);
    const int64_t r0 = k0 + b, m = n - r0;
    const int kk = (int)std::min<int64_t>(m, b);
    double* P = H + r0 + k0 * ldh;
    double* Tall = ar.take<double>(256 * 256);
    // [V | Y | V] side by side (ld m): [V Y] and [Y V] are the two operands of
    // the rank-2b update
    double* V = ar.take<double>((size_t)m * 3 * b);
    double* Y = V + (size_t)m * b;
    double* V2 = Y + (size_t)m * b;
    double* VT = ar.take<double>((size_t)m * b);
    double* M = ar.take<double>((size_t)b * b);
    double* Wm = ar.take<double>((size_t)b * b);
    ZOLO_ARENA_OK(ar);
    ZOLO_CHECK(dense_geqrf_cols<double>(c, P, ldh, m, kk, ts.tau1 + k0, Tall, ar));
    ts.ncols1 = k0 + kk;
    ZOLO_CHECK(ke_trd_extract_v(H, ldh, k0, r0, kk, m, V, m, b, c.s));
    if (kk < b) {
      // fewer rows than columns (the last panel): Q^T on the panel's other columns
      GemmDesc d;  // Wm = V^T P2
      d.transA = 1; d.M = kk; d.N = b - kk; d.K = m;
      d.A = V; d.lda = m; d.B = P + kk * ldh; d.ldb = ldh; d.C = Wm; d.ldc = kk;
      ZOLO_CHECK(gemm(Prec::F64, d, c.s));
      d = GemmDesc();  // M = T^T Wm
      d.transA = 1; d.M = kk; d.N = b - kk; d.K = kk;
      d.A = Tall; d.lda = 256; d.B = Wm; d.ldb = kk; d.C = M; d.ldc = kk; d.aHi = 0;
      ZOLO_CHECK(gemm(Prec::F64, d, c.s));
      d = GemmDesc();  // P2 -= V M
      d.M = m; d.N = b - kk; d.K = kk;
      d.A = V; d.lda = m; d.B = M; d.ldb = kk; d.C = P + kk * ldh; d.ldc = ldh; d.alpha = -1.0; d.beta = 1.0;
      ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    }
    double* A22 = H + r0 + r0 * ldh;
    GemmDesc d;  // VT = V T   (T upper: op(B)[k][j] = 0 for k > j)
    d.M = m; d.N = kk; d.K = kk;
    d.A = V; d.lda = m; d.B = Tall; d.ldb = 256; d.C = VT; d.ldc = m; d.bHi = 0;
    ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    d = GemmDesc();  // Y = A22 VT
    d.M = m; d.N = kk; d.K = m;
    d.A = A22; d.lda = ldh; d.B = VT; d.ldb = m; d.C = Y; d.ldc = m;
    ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    d = GemmDesc();  // M = V^T Y
    d.transA = 1; d.M = kk; d.N = kk; d.K = m;
    d.A = V; d.lda = m; d.B = Y; d.ldb = m; d.C = M; d.ldc = kk;
    ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    d = GemmDesc();  // Wm = T^T M   (T^T lower: op(A)[i][k] = 0 for k > i)
    d.transA = 1; d.M = kk; d.N = kk; d.K = kk;
    d.A = Tall; d.lda = 256; d.B = M; d.ldb = kk; d.C = Wm; d.ldc = kk; d.aHi = 0;
    ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    d = GemmDesc();  // Y -= V Wm / 2
    d.M = m; d.N = kk; d.K = kk;
    d.A = V; d.lda = m; d.B = Wm; d.ldb = kk; d.C = Y; d.ldc = m; d.alpha = -0.5; d.beta = 1.0;
    ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    if (kk == b) {
      // A22 -= [V Y] [Y V]^T = V Y^T + Y V^T: one pass over A22 with K = 2b
      ZOLO_HIP(hipMemcpyAsync(V2, V, sizeof(double) * (size_t)m * b, hipMemcpyDeviceToDevice, c.s));
      d = GemmDesc();
      d.transB = 1; d.M = m; d.N = m; d.K = 2 * b;
      d.A = V; d.lda = m; d.B = Y; d.ldb = m; d.C = A22; d.ldc = ldh; d.alpha = -1.0; d.beta = 1.0;
      ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    } else {
      d = GemmDesc();  // A22 -= V Y^T
      d.transB = 1; d.M = m; d.N = m; d.K = kk;
      d.A = V; d.lda = m; d.B = Y; d.ldb = m; d.C = A22; d.ldc = ldh; d.alpha = -1.0; d.beta = 1.0;
      ZOLO_CHECK(gemm(Prec::F64, d, c.s));
      d.A = Y; d.B = V;  // A22 -= Y V^T
      ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    }
  }
  ar.reset();
  ar.take<char>(mark);
  return kOk;
}

// Band -> tridiagonal (bulge chasing on the GPU), d and e out.
int sb2st(const Ctx& c, const double* H, int64_t ldh, int64_t n, double* d, double* e, TwoStage& ts) {
  const int b = kSbNb, g = kSbGroup;
  ZOLO_CHECK(ke_band_extract(H, ldh, n, b, ts.LB, ts.ldb, c.s));
  ZOLO_HIP(hipMemsetAsync(ts.progress, 0, sizeof(int) * ((size_t)n + 64), c.s));
  ZOLO_HIP(hipMemsetAsync(ts.vstore, 0, sizeof(double) * (size_t)ts.ngroups * g * ts.span, c.s));
  ZOLO_HIP(hipMemsetAsync(ts.tstore, 0, sizeof(double) * (size_t)ts.ngroups * g, c.s));
  // concurrent sweeps: sweep i+1 trails sweep i by three tasks
  const int nwg = (int)std::min<int64_t>(256, std::max<int64_t>(8, (n / b) / 3 + 8));
  ZOLO_CHECK(ke_sb2st(ts.LB, ts.ldb, n, g, ts.kmax, ts.vstore, ts.tstore, ts.progress, ts.status, nwg, c.s));
  int st = 0;
  ZOLO_HIP(hipMemcpyAsync(&st, ts.status, sizeof(int), hipMemcpyDeviceToHost, c.s));
  ZOLO_HIP(hipStreamSynchronize(c.s));
  if (st != 0) {
    set_error("sb2st: a sweep waited longer than its spin budget");
    return kErrInternal;
  }
  return ke_band_diag(ts.LB, ts.ldb, n, d, e, c.s);
}

// Z <- Q1 Q2 Z (Z n x n, ld ldz): the bulge-chasing reflectors in block
// reflectors of g sweeps at one task index, in anti-diagonal wavefronts of
// mutually disjoint groups (one batched launch per GEMM and wavefront), then
// the stage-1 reflectors in blocks of 256 columns, last block first.
int back_two_stage(const Ctx& c, const double* H, int64_t ldh, int64_t n, double* Z, int64_t ldz, TwoStage& ts,
                   Arena& ar) {
  const int b = kSbNb, g = kSbGroup;
  const int64_t ldp = ts.ldzp, span = ts.span;
  // padded copy: the group row ranges reach past row n - 1 with zero V rows
  ZOLO_CHECK(ke_fill(ts.Zp, ldp, ldp, n, 0.0, c.s));
  ZOLO_HIP(hipMemcpy2DAsync(ts.Zp, sizeof(double) * ldp, Z, sizeof(double) * ldz, sizeof(double) * n, n,
                            hipMemcpyDeviceToDevice, c.s));
  if (ts.ngroups > 0 && n >= 3) {
    GemmDesc d;  // G = V^T V per group
    d.batch = ts.ngroups;
    d.transA = 1; d.M = g; d.N = g; d.K = span;
    d.A = ts.vstore; d.lda = span; d.strideA = (int64_t)g * span;
    d.B = ts.vstore; d.ldb = span; d.strideB = (int64_t)g * span;
    d.C = ts.Gs; d.ldc = g; d.strideC = (int64_t)g * g;
    ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    ZOLO_CHECK(ke_q2_form_t(ts.Gs, ts.tstore, ts.ngroups, ts.Ts, c.s));
    const int J = ts.J, K = ts.kmax;
    for (int t = 0; t <= J + K - 2; ++t) {
      // groups (j, k = t - (J - 1 - j)) with 0 <= k < K and sweep j g of task k present
      const int64_t base_k = (int64_t)t - (J - 1);
      int64_t jlo = std::max<int64_t>(0, -base_k);
      int64_t jhi = std::min<int64_t>(J - 1, K - 1 - base_k);
      const int64_t lim = n - 2 - base_k * b;  // j (g + b) <= n - 2 - base_k b
      if (lim < 0) continue;
      jhi = std::min<int64_t>(jhi, lim / (g + b));
      if (jhi < jlo) continue;
      const int cnt = (int)(jhi - jlo + 1);
      const int64_t k0 = base_k + jlo;
      const int64_t grp0 = jlo * K + k0;
      const int64_t row0 = jlo * g + 1 + k0 * b;
      const int64_t sgrp = K + 1;  // group index stride along the wavefront
      GemmDesc w;  // W = V^T Zp(rows)   (V(r, q) = 0 unless q <= r <= q + b - 1)
      w.batch = cnt;
      w.transA = 1; w.M = g; w.N = n; w.K = span;
      w.A = ts.vstore + grp0 * g * span; w.lda = span; w.strideA = sgrp * g * span;
      w.B = ts.Zp + row0; w.ldb = ldp; w.strideB = g + b;
      w.C = ts.W; w.ldc = g; w.strideC = (int64_t)g * n;
      w.aLo = 0; w.aHi = b - 1;
      ZOLO_CHECK(gemm(Prec::F64, w, c.s));
      GemmDesc t2;  // VT = V T (span x g per group, in W2; T upper)
      t2.batch = cnt;
      t2.M = span; t2.N = g; t2.K = g;
      t2.A = ts.vstore + grp0 * g * span; t2.lda = span; t2.strideA = sgrp * g * span;
      t2.B = ts.Ts + grp0 * g * g; t2.ldb = g; t2.strideB = sgrp * g * g;
      t2.C = ts.W2; t2.ldc = span; t2.strideC = span * g;
      t2.aLo = b - 1; t2.aHi = 0; t2.bHi = 0;
      ZOLO_CHECK(gemm(Prec::F64, t2, c.s));
      GemmDesc u;  // Zp(rows) -= (V T) W   (VT(r, q) = 0 for r > q + b - 1)
      u.batch = cnt;
      u.M = span; u.N = n; u.K = g;
      u.A = ts.W2; u.lda = span; u.strideA = span * g;
      u.B = ts.W; u.ldb = g; u.strideB = (int64_t)g * n;
      u.C = ts.Zp + row0; u.ldc = ldp; u.strideC = g + b;
      u.alpha = -1.0; u.beta = 1.0;
      u.aLo = b - 1;
      ZOLO_CHECK(gemm(Prec::F64, u, c.s));
    }
  }
  // stage 1
  const size_t mark = ar.used();
  for (int64_t P = ((ts.ncols1 - 1) / 256) * 256; ts.ncols1 > 0 && P >= 0; P -= 256) {
    const int jb = (int)std::min<int64_t>(256, ts.ncols1 - P);
    ar.reset();
    ar.take<char>(mark);
    ZOLO_CHECK(apply_trd_block(c, H, ldh, ts.tau1, P, jb, n, ts.Zp, ldp, n, ar, b));
  }
  ar.reset();
  ar.take<char>(mark);
  ZOLO_HIP(hipMemcpy2DAsync(Z, sizeof(double) * ldz, ts.Zp, sizeof(double) * ldp, sizeof(double) * n, n,
                            hipMemcpyDeviceToDevice, c.s));
  return kOk;
}

int take_two_stage(Arena& ar, int64_t n, TwoStage& ts) {
  const int64_t b = kSbNb, g = kSbGroup;
  ts.n = n;
  ts.ldb = 2 * b;
  ts.span = b + g - 1;
  ts.J = (int)((n - 2 + g - 1) / g);
  ts.kmax = (int)((n - 2) / b + 1);
  ts.ngroups = ts.J * ts.kmax;
  ts.ldzp = n + ts.span + b + g;
  ts.tau1 = ar.take<double>(n);
  ts.LB = ar.take<double>((size_t)n * ts.ldb);
  ts.vstore = ar.take<double>((size_t)ts.ngroups * g * ts.span);
  ts.tstore = ar.take<double>((size_t)ts.ngroups * g);
  ts.Gs = ar.take<double>((size_t)ts.ngroups * g * g);
  ts.Ts = ar.take<double>((size_t)ts.ngroups * g * g);
  ts.progress = ar.take<int>((size_t)n + 64);
  ts.status = ts.progress + n;
  ts.Zp = ar.take<double>((size_t)ts.ldzp * n);
  ts.W = ar.take<double>((size_t)ts.J * g * n);
  ts.W2 = ar.take<double>((size_t)ts.J * g * n);
  ZOLO_ARENA_OK(ar);
  return kOk;
}

}  // namespace

size_t syevd_bytes(int64_t n) {
  const size_t nn = (size_t)n * n;
  const size_t chunks = (size_t)((n + 63) / 64);  // symmetric matrix-vector partial slots
  size_t b = 0;
  b += sizeof(double) * ((size_t)n * kTrdNb + 5 * (size_t)n + 64 + 8 + 64 * 2 * kTrdNb + 64 + chunks * n);  // trd
  b += sizeof(double) * (2 * nn);                                       // Qn, U
  b += sizeof(double) * (14 * (size_t)n);                               // D&C double arrays
  b += sizeof(int) * (10 * (size_t)n);                                  // D&C int arrays
  b += (sizeof(MergeH) + 3 * sizeof(int) + sizeof(double)) * (size_t)n;  // merges, counts, rho
  const size_t nbo = 256;
  b += sizeof(double) * ((size_t)n * nbo + 3 * nbo * nbo + 2 * nbo * (size_t)n);  // back-transformation
  if (n >= kTwoStageMin) b += two_stage_bytes(n);
  return b + 64 * 256;
}

int syevd_run(int64_t n, double* H, int64_t ldh, double* lambda, double* Z, int64_t ldz, hipStream_t s,
              double* phase_ms) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const size_t bytes = syevd_bytes(n);
  char* ws = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &ws));
  Arena ar(ws, bytes);
  const int64_t nn = n * n;
  Ctx c;
  c.s = s;
  // ---- 1. tridiagonalisation --------------------------------------------------
  double* W = ar.take<double>((size_t)n * kTrdNb);
  double* v = ar.take<double>(n);
  double* tau = ar.take<double>(n);
  double* d = ar.take<double>(n);
  double* e = ar.take<double>(n);
  TrdScratch tw;
  tw.part = ar.take<double>(64);
  tw.sc = ar.take<double>(8);
  tw.dpart = ar.take<double>(64 * 2 * kTrdNb);
  tw.wpart = ar.take<double>(64);
  tw.ypart = ar.take<double>((size_t)((n + 63) / 64) * n);
  ZOLO_ARENA_OK(ar);
  ZOLO_CHECK(ke_fill(W, n, n, kTrdNb, 0.0, s));
  ZOLO_CHECK(ke_fill(e, n, n, 1, 0.0, s));
  const bool two = n >= kTwoStageMin;
  TwoStage ts;
  if (two) {
    ZOLO_CHECK(take_two_stage(ar, n, ts));
    ZOLO_CHECK(sy2sb(c, H, ldh, n, ts, ar));
    ZOLO_CHECK(sb2st(c, H, ldh, n, d, e, ts));
  }
  for (int64_t k0 = 0; !two && k0 < n; k0 += kTrdNb) {
    const int jb = (int)std::min<int64_t>(kTrdNb, n - k0);
    for (int64_t i = k0; i < k0 + jb; ++i)
      ZOLO_CHECK(ke_trd_column(H, ldh, n, k0, i, W, v, tau, e, tw, i + 1 == k0 + jb, s));
    ZOLO_CHECK(ke_trd_panel_end(H, ldh, k0, jb, n, d, s));
    const int64_t r = k0 + jb, n2 = n - r;
    if (n2 <= 0) continue;
    GemmDesc g;  // A22 -= V W^T
    g.transB = 1; g.M = n2; g.N = n2; g.K = jb;
    g.A = H + r + k0 * ldh; g.lda = ldh;
    g.B = W + r; g.ldb = n;
    g.C = H + r + r * ldh; g.ldc = ldh;
    g.alpha = -1.0; g.beta = 1.0;
    ZOLO_CHECK(gemm(Prec::F64, g, s));
    g.A = W + r; g.lda = n;  // A22 -= W V^T
    g.B = H + r + k0 * ldh; g.ldb = ldh;
    ZOLO_CHECK(gemm(Prec::F64, g, s));
  }
  if (phase_ms) ZOLO_HIP(hipStreamSynchronize(s));
  const auto t1 = clk::now();
  // ---- 2. divide and conquer --------------------------------------------------
  const std::vector<Level> levels = plan_levels(n);
  size_t nm = 0;
  for (auto& L : levels) nm += L.merges.size();
  double* Qn = ar.take<double>(nn);
  double* U = ar.take<double>(nn);
  double* lamA = ar.take<double>(n);
  double* lamB = ar.take<double>(n);
  DcLevel D;
  D.dK = ar.take<double>(n);
  D.zK = ar.take<double>(n);
  D.dD = ar.take<double>(n);
  D.rotcs = ar.take<double>(2 * n);
  D.zbuf = ar.take<double>(n);
  D.dbuf = ar.take<double>(n);
  D.tau = ar.take<double>(n);
  D.zh = ar.take<double>(n);
  D.inv_norm = ar.take<double>(n);
  D.vals = ar.take<double>(n);
  D.pK = ar.take<int>(n);
  D.pD = ar.take<int>(n);
  D.rotab = ar.take<int>(2 * n);
  D.perm = ar.take<int>(n);
  D.org = ar.take<int>(n);
  D.rank = ar.take<int>(n);
  MergeH* dmerges = ar.take<MergeH>(std::max<size_t>(1, nm));
  D.cnt = ar.take<int>(3 * std::max<size_t>(1, nm));
  D.rho = ar.take<double>(std::max<size_t>(1, nm));
  ZOLO_ARENA_OK(ar);
  if (nm > 0) {
    std::vector<MergeH> all;
    for (auto& L : levels) all.insert(all.end(), L.merges.begin(), L.merges.end());
    ZOLO_HIP(hipMemcpyAsync(dmerges, all.data(), sizeof(MergeH) * all.size(), hipMemcpyHostToDevice, s));
  }
  double* Q = Z;
  int64_t ldQ = ldz;
  double* Qo = Qn;
  int64_t ldQo = n;
  ZOLO_CHECK(ke_fill(Q, ldQ, n, n, 0.0, s));
  ZOLO_CHECK(ke_fill(Qo, ldQo, n, n, 0.0, s));
  ZOLO_CHECK(ke_dc_leaves(d, e, n, lamA, Q, ldQ, s));
  double* lam = lamA;
  double* lam2 = lamB;
  for (const Level& L : levels) {
    int smax = 0;
    for (auto& m : L.merges) smax = std::max(smax, m.s1 + m.s2);
    ZOLO_CHECK(ke_fill(U, n, n, n, 0.0, s));
    D.nmerge = (int)L.merges.size();
    D.smax = smax;
    D.merges = dmerges + L.first;
    D.e = e;
    D.lam = lam;
    D.lam_new = lam2;
    D.Q = Q;
    D.ldq = ldQ;
    D.U = U;
    D.ldu = n;
    ZOLO_CHECK(ke_dc_level(D, s));
    // Q_new(block) = Q(block) U(block), merges of equal size in one strided batch
    size_t a = 0;
    while (a < L.merges.size()) {
      const int sz = L.merges[a].s1 + L.merges[a].s2;
      size_t b = a + 1;
      while (b < L.merges.size() && L.merges[b].s1 + L.merges[b].s2 == sz &&
             L.merges[b].off == L.merges[b - 1].off + sz)
        ++b;
      const int off = L.merges[a].off;
      GemmDesc g;
      g.batch = (int)(b - a);
      g.M = sz; g.N = sz; g.K = sz;
      g.A = Q + off + off * ldQ; g.lda = ldQ; g.strideA = (int64_t)sz * (ldQ + 1);
      g.B = U + off + off * n; g.ldb = n; g.strideB = (int64_t)sz * (n + 1);
      g.C = Qo + off + off * ldQo; g.ldc = ldQo; g.strideC = (int64_t)sz * (ldQo + 1);
      ZOLO_CHECK(gemm(Prec::F64, g, s));
      a = b;
    }
    if (L.carry_s > 0) {
      const int o = L.carry_off, cs = L.carry_s;
      ZOLO_HIP(hipMemcpy2DAsync(Qo + o + o * ldQo, sizeof(double) * ldQo, Q + o + o * ldQ, sizeof(double) * ldQ,
                                sizeof(double) * cs, cs, hipMemcpyDeviceToDevice, s));
      ZOLO_HIP(hipMemcpyAsync(lam2 + o, lam + o, sizeof(double) * cs, hipMemcpyDeviceToDevice, s));
    }
    std::swap(Q, Qo);
    std::swap(ldQ, ldQo);
    std::swap(lam, lam2);
  }
  if (Q != Z) ZOLO_HIP(hipMemcpy2DAsync(Z, sizeof(double) * ldz, Q, sizeof(double) * ldQ, sizeof(double) * n, n,
                                        hipMemcpyDeviceToDevice, s));
  ZOLO_HIP(hipMemcpyAsync(lambda, lam, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if (phase_ms) ZOLO_HIP(hipStreamSynchronize(s));
  const auto t2 = clk::now();
  // ---- 3. back-transformation: Z <- Q_trd Z ------------------------------------
  const size_t mark = ar.used();
  if (two) ZOLO_CHECK(back_two_stage(c, H, ldh, n, Z, ldz, ts, ar));
  const int64_t nref = two ? 0 : n - 1;  // reflectors k = 0 .. n-2
  for (int64_t P = ((nref - 1) / 256) * 256; nref > 0 && P >= 0; P -= 256) {
    const int jb = (int)std::min<int64_t>(256, nref - P);
    ar.reset();
    ar.take<char>(mark);
    ZOLO_CHECK(apply_trd_block(c, H, ldh, tau, P, jb, n, Z, ldz, n, ar));
  }
  if (phase_ms) {
    ZOLO_HIP(hipStreamSynchronize(s));
    const auto t3 = clk::now();
    phase_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    phase_ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    phase_ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
  }
  return kOk;
}

}  // namespace zolo
