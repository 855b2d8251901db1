/*
 * zolo.h -- C ABI of the MI355X (gfx950) Zolo-PD / QDWH-PD library (libzolo.so).
 *
 * All matrices are column-major FP64 with an explicit leading dimension, as on
 * a 1x1 ScaLAPACK grid of the reference.  Every entry point returns an int
 * status: 0 on success, a negative value -i when argument i is illegal
 * (ScaLAPACK convention), or one of the positive ZOLO_ERR_* codes below.  The
 * library never aborts; zolo_last_error() describes the last failure.
 *
 * Threading: calls are serialised on one device context (a grow-only device
 * workspace owned by the library); they are not re-entrant.  `stream` is a
 * hipStream_t (NULL = the default stream).
 */
#ifndef ZOLO_H_
#define ZOLO_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZOLO_OK 0
#define ZOLO_ERR_HIP 1001            /* HIP runtime error */
#define ZOLO_ERR_NOT_SPD 1002        /* a shifted Gram matrix was not positive definite */
#define ZOLO_ERR_NO_CONVERGENCE 1003 /* iteration cap reached above tolerance */
#define ZOLO_ERR_ALLOC 1004          /* device allocation failed */
#define ZOLO_ERR_INTERNAL 1005

/* Arithmetic of the Cholesky-based iterations. */
#define ZOLO_PREC_F64 0     /* FP64 throughout (the reference's arithmetic)               */
#define ZOLO_PREC_F32 1     /* FP32 Cholesky iterations, FP64 QR step + polish            */
#define ZOLO_PREC_BF16X3 2  /* BF16x3 MFMA Cholesky iterations, FP64 QR step + polish     */
/* The low-precision Cholesky steps above take the fewest shifts whose
 * Zolotarev bound reaches FP32 working accuracy in that step and stop once
 * CON - 1 is below the FP32 unit roundoff.  These two keep the reference's
 * schedule instead (all r shifts every step, the loop ends on the ERR test of
 * pzolopd1.f90:397), then polish in FP64 the same way; they separate what the
 * precision buys from what the shorter schedule buys. */
#define ZOLO_PREC_F32_REFSCHED 3
#define ZOLO_PREC_BF16X3_REFSCHED 4

/* Per-run record: iteration trajectory and phase times. */
typedef struct zolo_info {
  int iters;              /* iterations executed                                       */
  int r;                  /* Zolotarev degree (number of shifts)                       */
  int info;               /* first non-positive pivot column + 1 of a shifted Cholesky */
  int polish_passes;      /* FP64 polish passes executed (mixed precision only)         */
  int restarts;           /* PZOLOPD2 restarts: bounds re-estimated when CON reached 1
                             before ERR met TOL2 (pzolopd2.f90:397-405)                  */
  char schedule[16];      /* per iteration: 'Q' QR step, 'C' FP64 Cholesky, 'c' low prec*/
  double err[16];         /* ERR = ||X+ - X||_F after each FormX2                      */
  double con[16];         /* CON at the start of each iteration                        */
  double maxc[16];        /* max c_i of each iteration                                 */
  double alpha, beta;     /* bounds used                                               */
  double time_ms;         /* device time of the whole call (hipEvent)                  */
  double phase_ms[8];     /* 0 bounds, 1 QR steps, 2 Cholesky steps, 3 FormX2,
                             4 polish, 5 H, 6 other                                    */
  int shifts[16];         /* shifts used by each iteration: r, except a low-precision
                             Cholesky step, which takes the fewest shifts whose bound
                             reaches FP32 convergence in that step                     */
  int cg_iters[4];        /* conjugate-gradient iterations of each polish pass          */
} zolo_info_t;

/* ---- library lifetime ---------------------------------------------------- */
int zolo_init(int device);
int zolo_finalize(void);
const char* zolo_last_error(void);
int zolo_version(void);

/* ---- scalar helpers: Auxil_Zolo.f90:36-200, hfunc.f90:2-42 ---------------- */
/* coeff[0:2r] = c_1..c_2r, coeff[2r:3r] = a_1..a_r (Compute_Coeff, Auxil_Zolo.f90:165) */
int zolo_compute_coeff(double con, double tol, int rk, double* coeff);
int zolo_choosem(double con);                                          /* Auxil_Zolo.f90:130 */
double zolo_hfunc(double x);                                           /* hfunc.f90:2        */
void zolo_mellipke(double alpha, double tol, double* k, double* e);    /* Auxil_Zolo.f90:36  */
void zolo_mellipj(double u, double alpha, double tol, double* sn, double* cn, double* dn); /* :73 */

/* ---- Zolo-PD: replaces PZOLOPD1 (pzolopd1.f90:2) on a 1x1 grid ----------------
 * U = polar factor of A (n x n), H = (U^T A + A^T U)/2 if H != NULL.
 * sym = 'S' or 'N' (zolochol.f90:99), rk = number of shifts (1..8),
 * alpha, beta as in pzolopd1: X0 = A/(alpha*beta), CON = 1/beta.
 * Device pointers; A is read-only.  Every argument is checked before any
 * device work: -i for argument i (A NULL -7, lda < n -8, U NULL -9, ...). */
int zolo_pd_device(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A,
                   int64_t lda, double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info, void* stream);
/* zolo_pd_device with the bounds computed inside (no counterpart in the
 * reference, whose driver computes alpha, beta by LU + condition estimate,
 * test_zolodrv1.f90:179-189): alpha = ||A||_F, A/alpha = Q_X R by the same
 * Householder QR the first QR step of the iteration needs, and beta =
 * 1/(||R^{-1}||_1 estimate * sqrt(n)) with Higham's estimator (DTRCON's);
 * sigma_min(A/alpha) = 1/||R^{-1}||_2 >= beta as for the LU estimate.  The
 * first QR step then starts from that factorisation.  rk = 0 picks
 * choosem(1/beta); info->alpha, info->beta, info->r report the choice.  For
 * a well-conditioned A whose first step is a Cholesky step (largest shift
 * coefficient <= 300, pzolopd1.f90:405) the QR is not reused, and
 * zolo_bounds_device + zolo_pd_device is the cheaper route.
 * Errors by this function's own argument positions (precision -1, sym -2,
 * rk -3, n -4, A NULL -5, lda -6, U NULL -7, ldu -8, ldh -10), all checked
 * before any device work; a numerically singular A (no positive bound) is
 * kErrArg. */
int zolo_pd_auto_device(int precision, char sym, int rk, int64_t n, const double* A, int64_t lda, double* U,
                        int64_t ldu, double* H, int64_t ldh, zolo_info_t* info, void* stream);
/* Same with host arrays (copied to and from the device). */
int zolo_pd_host(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A,
                 int64_t lda, double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info);

/* Replaces PZOLOPD2 (pzolopd2.f90:2): as zolo_pd_host with the iteration cap of
 * pzolopd2.f90:394 (12 instead of 6).  The restart branch of pzolopd2 calls
 * FormXTwo, which the reference does not define; in its place, when CON has
 * reached 1 while ERR is still above TOL2, the bounds of the current iterate are
 * re-estimated and the iteration continues from it (info.restarts counts them). */
int zolo_pd2_host(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A,
                  int64_t lda, double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info);

/* ---- Bounds: alpha = ||A||_F, beta from LU + 1-norm estimate ---------------
 * test_zolodrv1.f90:179-189 (convention 0: beta = rcond/sqrt(n)) and
 * pqdwhfacs.f90:318-344 (convention 1: beta = ||A/alpha||_1 rcond/sqrt(n)). */
int zolo_bounds_device(int convention, int64_t n, const double* A, int64_t lda, double* alpha, double* beta,
                       void* stream);
int zolo_bounds_host(int convention, int64_t n, const double* A, int64_t lda, double* alpha, double* beta);

/* ---- QDWH-PD: replaces PQDWHFACS (pqdwhfacs.f90:2) on a 1x1 grid ---------- */
int zolo_qdwh_device(char sym, int64_t n, double sigma, const double* A, int64_t lda, double* U, int64_t ldu,
                     double* H, int64_t ldh, zolo_info_t* info, void* stream);
int zolo_qdwh_host(char sym, int64_t n, double sigma, const double* A, int64_t lda, double* U, int64_t ldu, double* H,
                   int64_t ldh, zolo_info_t* info);

/* ---- Single-shift steps: replace ZOLOCHOL (zolochol.f90:2) / ZOLOQR (zoloqr.f90:2)
 * Z = -a_ip X (X^T X + c_{2ip-1} I)^{-1}  and  Z = -a_ip/sqrt(c) Q1 Q2^T, ip 1-based.
 * coeff holds 3*rk values as from zolo_compute_coeff. Host arrays. */
int zolo_chol_step_host(char sym, int64_t n, const double* X, int64_t ldx, double* Z, int64_t ldz, const double* coeff,
                        int rk, int ip);
int zolo_qr_step_host(int64_t n, const double* X, int64_t ldx, double* Z, int64_t ldz, const double* coeff, int rk,
                      int ip);

/* ---- 2-norm estimate: replaces PDNORMEST (pdnormest.f90:2) ----------------- */
int zolo_normest_device(int64_t m, int64_t n, const double* A, int64_t lda, double tol, double* est, int* cnt,
                        void* stream);
int zolo_normest_host(int64_t m, int64_t n, const double* A, int64_t lda, double tol, double* est, int* cnt);
/* No counterpart in the reference.  zolo_normest_device for callers that
 * cannot allocate device memory (graph capture, serving loops, harnesses that
 * forbid hipMalloc): no allocation of any kind; the scratch (x, S x and the
 * partial sums, 8 (m + n + m ceil(n/256) + 1040) bytes + alignment) lives in a
 * fixed 16 MiB device array of the library, one per GPU, held until the
 * process ends.  Larger m, n return -1 with a message.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (other streams or host
 * threads) race; nothing in it needs to start at zero. */
int zolo_normest_noalloc_device(int64_t m, int64_t n, const double* A, int64_t lda, double tol, double* est, int* cnt,
                                void* stream);

/* ---- Structured QR of the 2n x n stack [X; s I] on host arrays -----------------
 * MPDGEQRF (mpdgeqrf.f:5): S <- R + reflectors, tau out.
 * MPDORGQR (mpdorgqr.f:5): S (reflectors) + tau -> S <- explicit Q (2n x n). */
int zolo_mpdgeqrf_host(int64_t n, double* S, int64_t lds, double* tau);
int zolo_mpdorgqr_host(int64_t n, double* S, int64_t lds, const double* tau);

/* ---- building blocks (device pointers; used by tests and the benchmark) --- */
/* C = alpha op(A) op(B) + beta C, precision 0 = F64, 1 = F32 (float data), 2 = BF16x3 (float data),
 * 3 = F64 data with the products formed exactly from int8 residues (16 moduli, int8 MFMA) and
 * the result rebuilt by the Chinese remainder theorem: FP64-level accuracy (inputs rounded to
 * 53 or more bits relative to their row / column maximum), K <= 131072 (larger K: FP64 MFMA) */
int zolo_gemm_device(int precision, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                     const void* A, int64_t lda, const void* B, int64_t ldb, double beta, void* C, int64_t ldc,
                     void* stream);
/* Structured QR of the 2n x n stack S = [X; s I] in place (R + reflectors), Q (2n x n) out, tau (n) out.
 * Restates MPDGEQRF + MPDORGQR (mpdgeqrf.f:5, mpdorgqr.f:5). */
int zolo_stacked_qr_device(int64_t n, double* S, int64_t lds, double* Q, int64_t ldq, double* tau, void* stream);
/* Lower Cholesky with explicit inverse: A (lower) -> L, Li = L^{-1}. */
int zolo_chol_inv_device(int precision, int64_t n, void* A, int64_t lda, void* Li, int64_t ldl, void* stream);

/* ---- batched small polar decompositions (FP32 data) -------------------------
 * count independent matrices A_i (n_i x n_i, 1 <= n_i <= 512, column-major FP32
 * in device memory, leading dimension lda_i) -> polar factors U_i (ldu_i).  Per
 * matrix this is PZOLOPD1 (pzolopd1.f90:280-487) on the bounds of
 * test_zolodrv1.f90:179-189 (alpha = ||A||_F, beta = rcond/sqrt(n) from an FP64
 * LU and the 1-norm estimator); the reference calls PZOLOPD1 once per matrix.
 * The iterations run in FP32 (precision must be 1: BF16x3 products leave the
 * orthogonality near 1e-5, above this path's 1e-6 target); the
 * schedule of every matrix follows from its beta, so all iterations of all
 * matrices run as one captured hipGraph.  A matrix stops after the step whose
 * CON - 1 is below the FP32 unit roundoff.
 * l0_source: 0 = the LU estimate, 1 = the learned estimate (zolo_l0_set_weights
 * first), 2 = beta_in[i] from the caller.  With 1 and 2, a matrix whose
 * ||U^T U - I||_F / sqrt(n_i) exceeds orth_tol falls back to the LU estimate:
 * it restarts from its iterate with the LU estimate of that iterate (below),
 * or, when the iterate is not finite, is recomputed from A with the LU
 * estimate of A.  rk > 0 overrides choosem.  tile: -1 autotune per bucket (cached
 * per process), 0 automatic, 1..4 fixed (128x128, 64x128, 128x64, 64x64).
 * use_graph = 0 launches eagerly.  Per-matrix outputs (host arrays, each may be
 * NULL): beta_out, orth_out, iters_out, flags_out (bit 0: fallback used,
 * bit 1: check failed or no positive bound, bit 2: restarted).  A matrix
 * still off the orthogonality target after its planned steps (the CON
 * recurrence assumes exact arithmetic; near CON ~ 1/u the first step is not
 * that accurate) restarts from its current iterate, which has the same polar
 * factor, with the LU bound of that iterate (as the restart of
 * pzolopd2.f90:397-405), at most twice. */
typedef struct zolo_batch_info {
  int count;             /* matrices                                              */
  int buckets;           /* size buckets (n rounded up to a multiple of 32)       */
  int fallbacks;         /* matrices whose given bound failed (LU estimate used)  */
  int failures;          /* matrices failing the orthogonality check at the end   */
  int max_iters;         /* largest iteration count                               */
  int graph_nodes;       /* nodes of the captured graphs                          */
  int continued;         /* matrices restarted from their iterate (see below)     */
  double time_ms;        /* wall time of the call                                 */
  double phase_ms[8];    /* 0 alpha + beta, 1 plan, 2 tile choice, 3 iterations,
                            4 fallback and restarts, 5 learned features + network */
  double flops_f32;      /* algorithmic flops of the executed schedules at each
                            matrix's own order n: per Cholesky step (7/3) n^3 per
                            shift + n^3, per FP32 QR step 5 n^3 per shift        */
  double flops_f64;      /* the same for the FP64 work: LU bounds (2/3) n^3 and
                            first QR steps at CON > 1e8 (5 n^3 per shift)        */
} zolo_batch_info_t;
int zolo_pd_batched_device(int count, const int64_t* n, const float* const* A, const int64_t* lda, float* const* U,
                           const int64_t* ldu, int precision, int l0_source, const double* beta_in, double orth_tol,
                           int rk, int tile, int use_graph, double* beta_out, double* orth_out, int* iters_out,
                           int* flags_out, zolo_batch_info_t* info, void* stream);
/* GEMM tile the autotuner chose for matrices of order n (-1 if not tuned yet). */
int zolo_batched_tile(int64_t n, int precision);

/* ---- Zolo-SVD tail (device pointers, FP64) ----------------------------------
 * zolo_syevd_device replaces ELPA_PDSYEVD2 (elpa_pdsyevd2.F90:1, called at
 * ZoloCHK1.f90:86): H (n x n symmetric, overwritten) = Z diag(w) Z^T, w
 * ascending.  Householder tridiagonalisation (MFMA rank-2k updates), Cuppen
 * divide and conquer on the tridiagonal (Gu-Eisenstat eigenvectors, MFMA
 * eigenvector updates), blocked back-transformation.  phase_ms (3 values,
 * may be NULL): tridiagonalisation, divide and conquer, back-transformation.
 * zolo_svd_device is the flow of ZoloCHK1.f90:80-91 on one GPU: polar factor
 * U_p by Zolo-PD with its bounds from the QR its first step reuses (as
 * zolo_pd_auto_device; precision as zolo_pd_device; rk = 0 picks choosem;
 * info->alpha/beta report them), H = (U_p^T A + A^T U_p)/2,
 * H = V diag(S) V^T (S ascending), U = U_p V; A = U diag(S) V^T.
 * phase_ms (5 values, may be NULL): polar + H, tridiagonalisation, divide and
 * conquer, back-transformation, eigensolve + U. */
int zolo_syevd_device(int64_t n, double* H, int64_t ldh, double* w, double* Z, int64_t ldz, double* phase_ms,
                      void* stream);
int zolo_svd_device(int precision, int rk, int64_t n, const double* A, int64_t lda, double* S, double* U, int64_t ldu,
                    double* V, int64_t ldv, zolo_info_t* info, double* phase_ms, void* stream);
/* zolo_svd_ex_device: zolo_svd_device with the choice of bounds (argument 2):
 * 0 = from the QR the first step reuses (what zolo_svd_device does), 1 = LU +
 * 1-norm estimate, alpha = ||A||_F and beta = rcond/sqrt(n), as the
 * reference's driver computes them before PZOLOPD1 (test_zolodrv1.f90:179-189;
 * ZoloCHK1 receives them).  The other arguments follow zolo_svd_device, one
 * position later. */
int zolo_svd_ex_device(int precision, int bounds, int rk, int64_t n, const double* A, int64_t lda, double* S,
                       double* U, int64_t ldu, double* V, int64_t ldv, zolo_info_t* info, double* phase_ms,
                       void* stream);
/* ---- spectral division (QDWH-eig / Zolo-eig; device pointers, FP64) ----------
 * zolo_subit_device replaces PQDWHSUBIT (pqdwhsubit.f90:2): on entry Q is the
 * polar factor of A - sigma I (A symmetric n x n, normA = ||A||_F > 0); C =
 * (Q - I)/2, Rk = round(||C||_F^2) = the number of eigenvalues below sigma;
 * subspace iteration (oversampling 3, at most two iterations) with full QR
 * gives V = [V1 V2].  On exit Q = V, A = V^T A V, *rk = Rk, *err (may be NULL)
 * = ||(V^T A V)(1:Rk, Rk+1:n)||_F / normA, *iters (may be NULL) = iterations;
 * the reference accepts err <= 4 * 10 eps / 2.
 * zolo_eig_device replaces ZoloEig (ZoloEig.f90:2): Q the polar factor of
 * As - sigma I, As the matrix (both overwritten); on exit w(1:Rk) the
 * eigenvalues below sigma, ascending, and V (n x Rk) their eigenvectors.
 * zolo_eig_below_device is the driver pqdwheig.f90:113-144 with the Zolo
 * polar factor: the polar factor of A - sigma I (precision and rk as
 * zolo_pd_auto_device, sym = 'S'), then zolo_eig_device; A is not modified,
 * *count = Rk, info (may be NULL) describes the polar decomposition. */
int zolo_subit_device(int64_t n, double* A, int64_t lda, double normA, double* Q, int64_t ldq, int* rk, double* err,
                      int* iters, void* stream);
int zolo_eig_device(int64_t n, double* Q, int64_t ldq, double* As, int64_t lda, double* w, double* V, int64_t ldv,
                    int* rk, double* err, int* iters, void* stream);
int zolo_eig_below_device(int precision, int rk, int64_t n, const double* A, int64_t lda, double sigma, double* w,
                          double* V, int64_t ldv, int* count, zolo_info_t* info, double* err, int* iters,
                          void* stream);
/* zolo_eig_below_ex_device: zolo_eig_below_device with the choice of bounds
 * (argument 2): 0 = from the QR the first step reuses, 1 = LU + 1-norm
 * estimate of A - sigma I with PQDWHFACS's convention, beta =
 * ||(A - sigma I)/alpha||_1 rcond/sqrt(n) (pqdwhfacs.f90:318-344, the bound of
 * the reference's pqdwheig flow).  The other arguments follow
 * zolo_eig_below_device, one position later. */
int zolo_eig_below_ex_device(int precision, int bounds, int rk, int64_t n, const double* A, int64_t lda, double sigma,
                             double* w, double* V, int64_t ldv, int* count, zolo_info_t* info, double* err,
                             int* iters, void* stream);
/* Host-array forms (the Fortran PQDWHSUBIT and ZoloEig drop-ins call them):
 * zolo_subit_host overwrites A and Q as zolo_subit_device; zolo_eig_host leaves
 * Q and As unchanged and returns w(1:rk) and V(:, 1:rk). */
int zolo_subit_host(int64_t n, double* A, int64_t lda, double normA, double* Q, int64_t ldq, int* rk, double* err,
                    int* iters);
int zolo_eig_host(int64_t n, double* Q, int64_t ldq, double* As, int64_t lda, double* w, double* V, int64_t ldv,
                  int* rk);
/* ---- MatrixMarket input (prepare_matrix.f90:154-321 prepare_real_GEM,
 * :638-759 prepare_real_symm_MM, mmread :1076) ---------------------------------
 * Coordinate files with field real, integer or pattern (value 1); complex and
 * the dense 'array' form are refused.  zolo_mm_info: order, entries, mtype (0
 * symmetric, 1 general, 2 skew-symmetric or hermitian, as prepare_real_GEM's
 * mtype) and, when bandwidth is not NULL, max |i - j| + 1 over the entries
 * (this reads the whole file).  zolo_mm_dense_host / _device: the square
 * matrix into A (column-major, lda >= n; host or device memory), zeroed first,
 * every entry stored in file order (a repeated position keeps its last value);
 * symmetric files then A <- A + A^T and A(i, i) <- A(i, i)/2 - sigma; other
 * types are left as stored (no shift).  The device form uploads the triplets
 * and assembles on the GPU (same values bit for bit).  Returns 0, or -1
 * cannot open, -2 banner/format, -3 size line, -4 entry line or early end,
 * -5 index outside the matrix, -6 not square, -7 A NULL, -8 lda < n. */
int zolo_mm_info(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz, int* mtype, int64_t* bandwidth);
int zolo_mm_dense_host(const char* path, double sigma, double* A, int64_t lda, int* mtype);
int zolo_mm_dense_device(const char* path, double sigma, double* A, int64_t lda, int* mtype, void* stream);
/* FormX2 (FormX2.f90:2) on host arrays: A <- (A + Z)/fone, Z <- the same, err = ||A_new - A_old||_F. */
int zolo_formx2_host(int64_t n, double fone, double* A, int64_t lda, double* Z, int64_t ldz, double* err);
/* Host-array form of zolo_syevd_device (the Fortran ELPA_PDSYEVD2 drop-in calls it). */
int zolo_syevd_host(int64_t n, double* H, int64_t ldh, double* w, double* Z, int64_t ldz);

/* ---- learned lower bound (replaces the condition estimate of
 * test_zolodrv1.f90:185-189 for the batched path, with it as the fallback) ----
 * Features of X = A/||A||_F (16 per matrix: log2 n, log10 of 8 Lanczos Ritz
 * values of X^T X, log10 ||X||_1, log10 ||X||_inf, log10 |r_kk| of the R factor
 * of a 4-column Gaussian sketch X Omega, n ||X||_2^2 estimate) feed a
 * 16 -> 64 -> 64 -> 1 tanh network predicting y = log10(sigma_min/sigma_max);
 * beta = 10^(y - margin) * sqrt(theta_1).  Parameters (float, in order):
 * feature mean[16], feature scale[16], W1[64][16], b1[64], W2[64][64], b2[64],
 * w3[64], b3, margin -- zolo_l0_param_count() values.  Training minimises the
 * pinball (quantile) loss mean rho_tau(t - y) with Adam, full batch, on the
 * device.  feat is count x 16 (row per matrix), all device pointers unless
 * noted. */
int zolo_l0_param_count(void);
int zolo_l0_set_weights(const float* params /* host */, int nparams);
int zolo_l0_features_device(int count, const int64_t* n, const float* const* A, const int64_t* lda, float* feat,
                            float* smax, void* stream);
int zolo_l0_forward_device(const float* params, const float* feat, int N, float* y, void* stream);
/* gradient of the loss w.r.t. the trained block (W1 .. b3), loss to a host float */
int zolo_l0_grad_device(const float* params, const float* feat, const float* target, int N, float tau, float* grad,
                        float* loss, void* stream);
/* steps of Adam (lr, beta1 0.9, beta2 0.999) on params in place; loss_hist: host, steps values or NULL */
int zolo_l0_train_device(float* params, const float* feat, const float* target, int N, int steps, float lr, float tau,
                         float* loss_hist, void* stream);

/* ---- FP64 products of the mixed precisions (no counterpart in the reference) ----
 * The FP64 parts of precisions 1-4 (the sum over shifts sum_q alpha_q Q1_q Q2_q^T of the QR
 * step, the polish's U^T U and U^T X, H = (U^T A + A^T U)/2) take their large products
 * from exact int8 products modulo 16 moduli (mode 1, the default; see zolo_gemm_device
 * precision 3) or from FP64 MFMA (mode 0).  Precision 0 always uses FP64 MFMA.  The mode is
 * process-wide; it returns -1 for a mode other than 0 or 1. */
int zolo_set_f64_products(int mode);
int zolo_get_f64_products(void);

/* ---- QR panel exchange (no counterpart in the reference; a test and tuning hook) ----
 * The panels of up to 8 shifts of at most 32 workgroups each run one shift per
 * XCD and, when every workgroup of a shift is found on one XCD, exchange their
 * column sums through that XCD's L2.  Mode 1 makes them use device-scope
 * stores instead (the form they fall back to when the placement check fails);
 * mode 0 (the default) is the automatic choice.  The results are the same
 * bits either way.  Process-wide; -1 for a mode other than 0 or 1. */
int zolo_set_panel_exchange(int mode);
int zolo_get_panel_exchange(void);

/* ---- measurement: GEMM launch ledger (no counterpart in the reference) ----
 * While enabled, every MFMA GEMM launch is bracketed by hipEvents on its stream
 * and its algorithmic flop count is recorded; zolo_profile_gemm sums launches
 * matching (precision, transa, transb, tile_m, tile_n), -1 = any.  Enabling
 * clears the ledger. */
int zolo_profile_enable(int on);
int zolo_profile_gemm(int precision, int transa, int transb, int tile_m, int tile_n, int64_t* launches, double* ms,
                      double* flops);

#ifdef __cplusplus
}
#endif
#endif /* ZOLO_H_ */
