// Symmetric eigensolver (tridiagonalisation, divide and conquer,
// back-transformation) and the Zolo-SVD assembly.
#pragma once
#include "linalg.hpp"

namespace zolo {

constexpr int kTrdNb = 32;  // tridiagonalisation panel width

struct TrdScratch {
  double* part;   // per-workgroup partial sums (<= 64)
  double* sc;     // device scalars
  double* dpart;  // per-workgroup partial dots (64 x 2 kTrdNb)
  double* ypart;  // matrix-vector partials (chunks x n)
  double* wpart;  // per-workgroup partial dots (<= 64)
};

struct DcLevel {
  int nmerge = 0;
  int smax = 0;
  const void* merges = nullptr;  // device array of {off, s1, s2}
  const double* e = nullptr;
  const double* lam = nullptr;
  double* lam_new = nullptr;
  double* Q = nullptr;
  int64_t ldq = 0;
  double* U = nullptr;
  int64_t ldu = 0;
  double *rho, *dK, *zK, *dD, *rotcs, *zbuf, *dbuf, *tau, *zh, *inv_norm, *vals;
  int *pK, *pD, *rotab, *cnt, *perm, *org, *rank;
};

int ke_fill(double* A, int64_t ld, int64_t m, int64_t n, double val, hipStream_t s);
// last_in_panel: the column applies its own final W correction; otherwise the
// next column's launch does (the columns of one panel run in order)
int ke_trd_column(double* A, int64_t ld, int64_t n, int64_t k0, int64_t i, double* W, double* v, double* tau, double* e,
                  TrdScratch& w, bool last_in_panel, hipStream_t s);
int ke_trd_panel_end(double* A, int64_t ld, int64_t k0, int jb, int64_t n, double* d, hipStream_t s);
// V (mrows x jb, ld ldv) = the reflectors of columns k0 .. k0+jb-1 of A, rows
// r0 .. r0+mrows-1: column c has its unit head at row k0 + c + off and its
// tail below it in A (off = 1 after the tridiagonalisation, the bandwidth
// after the full -> band stage); zeros above the head.
int ke_trd_extract_v(const double* A, int64_t ld, int64_t k0, int64_t r0, int jb, int64_t mrows, double* V,
                     int64_t ldv, int off, hipStream_t s);
int ke_dc_leaves(const double* d, const double* e, int64_t n, double* lam, double* Q, int64_t ldq, hipStream_t s);
int ke_dc_level(const DcLevel& L, hipStream_t s);

// Host drivers (eig_driver.cpp).
size_t syevd_bytes(int64_t n);
// H (n x n, ld) symmetric, overwritten; lambda ascending (n), Z eigenvectors (ld ldz).
int syevd_run(int64_t n, double* H, int64_t ldh, double* lambda, double* Z, int64_t ldz, hipStream_t s,
              double* phase_ms = nullptr);

// Spectral division (eig_subit.cpp).  subit_run: PQDWHSUBIT -- on entry Q is
// the polar factor of A - sigma I, on exit Q = V = [V1 V2] (n x n orthogonal,
// V1 spanning the eigenvectors of the rk eigenvalues below sigma) and A =
// V^T A V; err = ||E(1:rk, rk+1:n)||_F / normA, it = iterations (1 or 2).
int subit_run(int64_t n, double* A, int64_t lda, double normA, double* Q, int64_t ldq, int* rk, double* err, int* it,
              hipStream_t s);
// zoloeig_run: ZoloEig -- subit_run, then w (rk eigenvalues below sigma,
// ascending) and V (n x rk) their eigenvectors; Q and As are overwritten.
int zoloeig_run(int64_t n, double* Q, int64_t ldq, double* As, int64_t lda, double* w, double* V, int64_t ldv,
                int* rk, double* err, int* it, hipStream_t s);

// Apply Q = H_{k0} ... H_{k0+jb-1} (reflectors stored in A's columns k0..,
// heads at row k + off, tau[k]) from the left to C (rows k0+off .. n-1, ncols
// columns): C <- Q C.  Used by the back-transformations (off = 1 for the
// one-stage reduction, the bandwidth for the full -> band stage).
int apply_trd_block(const Ctx& c, const double* A, int64_t lda, const double* tau, int64_t k0, int jb, int64_t n,
                    double* C, int64_t ldc, int64_t ncols, Arena& ar, int off = 1);

// ---- two-stage reduction (sbr.hip) ----------------------------------------
constexpr int kSbNb = 64;     // bandwidth of the full -> band stage
constexpr int kSbGroup = 64;  // sweeps per block reflector of the bulge-chasing back-transformation
// LB (ld ldb >= 2b, n columns) <- the lower band d <= b of H, zero beyond
int ke_band_extract(const double* H, int64_t ldh, int64_t n, int b, double* LB, int64_t ldb, hipStream_t s);
// d = diagonal, e = subdiagonal of the band store (e[n-1] = 0)
int ke_band_diag(const double* LB, int64_t ldb, int64_t n, double* d, double* e, hipStream_t s);
// Bulge chasing of the band store (bandwidth kSbNb) to tridiagonal form on nwg
// workgroups; reflectors go to vstore / tstore in group layout (groups of g =
// kSbGroup sweeps, kmax tasks per sweep at most), progress (n ints) and
// status (1 int) must be zero on entry; status is 1 afterwards if a wait ran
// out of its spin budget.
int ke_sb2st(double* LB, int64_t ldb, int64_t n, int g, int kmax, double* vstore, double* tstore, int* progress,
             int* status, int nwg, hipStream_t s);
// T factors of the group block reflectors from G = V^T V (g x g per group)
int ke_q2_form_t(const double* Gstore, const double* tstore, int ngroups, double* Tstore, hipStream_t s);

}  // namespace zolo
