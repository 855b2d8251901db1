// Internal driver entry points behind the C ABI.
#pragma once
#include "../../include/zolo.h"
#include "linalg.hpp"

namespace zolo {

int zolo_pd_run(int precision, char sym, int rk, int64_t n, double alpha, double beta, const double* A, int64_t lda,
                double* U, int64_t ldu, double* H, int64_t ldh, zolo_info_t* info, hipStream_t s,
                int max_it = 6, bool restart = false);
// est = Higham's estimate of ||R^{-1}||_1 for the upper triangle of R (n x n,
// ld ldr; what lies below the diagonal is ignored), as DTRCON('1', 'U', 'N')
size_t rbound_bytes(int64_t n);
int rbound_run(int64_t n, const double* R, int64_t ldr, double* est, Arena& ar, hipStream_t s);
// MatrixMarket input (mm_reader.cpp)
int mm_info(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz, int* mtype, int64_t* bandwidth);
int mm_dense_host(const char* path, double sigma, double* A, int64_t lda, int* mtype);
int mm_dense_device(const char* path, double sigma, double* A, int64_t lda, int* mtype, hipStream_t s);
int qdwh_run(char sym, int64_t n, double sigma, const double* A, int64_t lda, double* U, int64_t ldu, double* H,
             int64_t ldh, zolo_info_t* info, hipStream_t s);
int form_h(const Ctx& c, const double* U, int64_t ldu, const double* A, int64_t lda, int64_t n, double* H,
           int64_t ldh);

// Bounds (test_zolodrv1.f90:179-189 / pqdwhfacs.f90:318-344): alpha = ||A||_F,
// beta from a GPU LU with partial pivoting and the Hager-Higham 1-norm estimator.
size_t bounds_bytes(int64_t n);
int bounds_run(int convention, int64_t n, const double* A, int64_t lda, double* alpha, double* beta, Arena& ar,
               hipStream_t s);
int normest_run(int64_t m, int64_t n, const double* A, int64_t lda, double tol, double* est, int* cnt, Arena& ar,
                hipStream_t s);
size_t normest_bytes(int64_t m, int64_t n);

// FP64 polish of a low-precision polar factor against an FP64 iterate Xref
// that has the same polar factor (first-order Lyapunov correction passes).
size_t polish_bytes(int64_t n);
// cg (4 entries, may be null): CG iterations of each pass.
int polish(const Ctx& c, double* U, int64_t ldu, const double* Xref, int64_t ldx, int64_t n, char* ws, int* passes,
           int* cg = nullptr);

}  // namespace zolo
