// Batched polar decomposition of many small FP32 matrices of uneven size.
#pragma once
#include <vector>

#include "linalg.hpp"

namespace zolo {

constexpr int kMaxBatchN = 512;  // largest matrix order of the batched path
constexpr int kBatchLuNb = 32;   // LU panel width (inverse diagonal block is 32 x 32)

// One matrix of a batch: n x n FP32, column-major (device pointers).
struct BatchMat {
  int64_t n;
  const float* A;
  int64_t lda;
  float* U;
  int64_t ldu;
};

// Kernels (batched.hip).  count = matrices of one bucket, items = (matrix,
// shift) pairs; idx maps a bucket slot to the global matrix index.
int kb_alpha(const BatchMat* mats, int count, double* alpha, hipStream_t s);
int kb_lu_pad(const BatchMat* mats, const int* idx, const double* alpha, double* W, int64_t nb, int count,
              hipStream_t s);
int kb_lu_panel(double* W, int64_t nb, int64_t k0, double* Linv, int* info, int count, hipStream_t s);
int kb_copy_block(const double* S, int64_t lds, int64_t sS, double* D, int64_t ldd, int64_t sD, int64_t m, int64_t n,
                  int count, hipStream_t s);
int kb_condest(const BatchMat* mats, const int* idx, const double* W, int64_t nb, const int* info, double* beta,
               int count, hipStream_t s);
template <typename T>
int kb_pad_scale(const BatchMat* mats, const int* idx, const double* alpha, const double* beta, T* X, int64_t nb,
                 int count, hipStream_t s);
template <typename T>
int kb_build_stack(const T* X, int64_t nb, const int* map, const double* sdiag, T* S, int items, hipStream_t s);
int kb_formx2_f64(const double* X64, const double* Z, float* X32, int64_t nb, const int* loc, const int* zoff,
                  const int* zcnt, const double* ifone, double* part, int nchunk, const int* gid, double* err2,
                  int count, hipStream_t s);
int kb_shift_copy(const float* W, int64_t nb, const int* map, const double* shift, float* Wq, int items,
                  hipStream_t s);
int kb_formx2(float* X, const float* Z, int64_t nb, const int* zoff, const int* zcnt, const double* ifone,
              double* part, int nchunk, const int* gid, double* err2, int count, hipStream_t s);
int kb_f2d(const float* X, double* Y, int64_t total, hipStream_t s);
int kb_gram_err(const double* G, int64_t nb, const BatchMat* mats, const int* idx, double* out, int count,
                hipStream_t s);
int kb_unpad(const float* X, int64_t nb, const BatchMat* mats, const int* idx, int count, hipStream_t s);
int kb_fill_int(int* p, int count, int value, hipStream_t s);

// Driver (batched_pd.cpp).
struct BatchOptions {
  int precision = 1;               // 1: FP32 iterations (the only one accepted)
  int l0_source = 0;               // 0: FP64 LU + 1-norm estimate; 2: beta_in (LU estimate as fallback)
  const double* beta_in = nullptr; // host, one per matrix
  double orth_tol = 1e-6;          // ||U^T U - I||_F / sqrt(n) accepted by the check
  int rk = 0;                      // Zolotarev degree override (0: choosem per matrix)
  int tile = -1;                   // GEMM tile: -1 autotune per bucket, 0 automatic, 1..4 fixed
  bool use_graph = true;           // capture each phase into one hipGraph
  int continue_rounds = 2;         // restarts of matrices still off after their planned steps
  bool one_bucket = false;         // all matrices padded to the largest bucket (few-matrix recomputations)
};

struct BatchResult {
  std::vector<double> alpha, beta, orth, err;
  std::vector<int> iters, rk, fallback, failed, continued;
  size_t graph_nodes = 0;
  int buckets = 0;
  double phase_ms[8] = {};         // 0 alpha + beta, 1 plan, 2 tile choice, 3 iterations, 4 fallback
  double time_ms = 0.0;
  // algorithmic flops of the executed work at each matrix's own order n (not
  // the padded bucket order): FP64 LU bound (2/3) n^3; per QR step 5 n^3 per
  // shift (FP64 for a first step at CON > 1e8); per Cholesky step (7/3) n^3
  // per shift + n^3 (the shared Gram matrix); fallback and restart calls added
  double flops_f32 = 0.0, flops_f64 = 0.0;
};

int pd_batched_run(int count, const int64_t* n, const float* const* A, const int64_t* lda, float* const* U,
                   const int64_t* ldu, const BatchOptions& opt, BatchResult& res, hipStream_t s);
// Tile the autotuner chose for a bucket of order nb (-1 if not tuned yet).
int batched_tile_of(int64_t nb, int precision);

// Learned lower bound (mlp.cpp): beta_i from the network on the matrix
// features; fails when no weights were set.
int l0_learned_beta(int count, const int64_t* n, const float* const* A, const int64_t* lda,
                    std::vector<double>& beta, double* ms, hipStream_t s);

}  // namespace zolo
