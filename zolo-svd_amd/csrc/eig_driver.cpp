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
  for (int64_t k0 = 0; k0 < n; k0 += kTrdNb) {
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
  const int64_t nref = n - 1;  // reflectors k = 0 .. n-2
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
