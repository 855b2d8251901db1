// Host-side launchers of the memory-bound / small kernels (kernels.hip, qr.hip).
#pragma once
#include "common.hpp"

namespace zolo {

constexpr int kCholLeaf = 64;   // diagonal block factored in LDS by one workgroup

template <typename T>
int k_formx2(T* X, int64_t ldx, const T* Z, int64_t ldz, int64_t strideZ, int nz, int64_t n, double fone,
             double* partial, int npartial, double* out, hipStream_t s);
template <typename T>
int k_sumsq(const T* A, int64_t lda, int64_t m, int64_t n, double* partial, int npartial, double* out, hipStream_t s);
template <typename T>
int k_norm1(const T* A, int64_t lda, int64_t m, int64_t n, double* colsum, double* out, hipStream_t s);
template <typename TA, typename TB>
int k_scale_copy(const TA* A, int64_t lda, TB* B, int64_t ldb, int64_t m, int64_t n, double alpha, hipStream_t s);
template <typename T>
int k_axpby(double a, const T* A, int64_t lda, double b, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t m,
            int64_t n, hipStream_t s);
template <typename T>
// S_q = [X; sdiag_q I] (upper: only the upper triangle of X, zeros below).
int k_build_stack(const T* X, int64_t ldx, T* S, int64_t lds, int64_t strideS, int64_t n, const double* sdiag,
                  int batch, hipStream_t s, bool upper = false);
template <typename T>
int k_shift_copy(const T* W, int64_t ldw, T* Wq, int64_t ldq, int64_t strideQ, int64_t n, const double* shift,
                 double scale, double add, int batch, hipStream_t s);
template <typename T>
int k_fill(T* A, int64_t lda, int64_t strideA, int64_t m, int64_t n, double value, int batch, hipStream_t s);
template <typename T>
int k_zero_upper(T* A, int64_t lda, int64_t strideA, int64_t n, int batch, hipStream_t s);
template <typename T>
int k_symmetrize(T* A, int64_t lda, int64_t n, hipStream_t s);
template <typename T>
int k_add_diag(T* A, int64_t lda, int64_t n, double d, hipStream_t s);
// X (m x n, ld) <- uniform [0, 1) numbers, a fixed function of (seed, index).
int k_fill_uniform(double* X, int64_t ld, int64_t m, int64_t n, unsigned long long seed, hipStream_t s);
// A(i, j) *= f on the upper triangle (i <= j) of an n x n block
int k_scale_upper(double* A, int64_t lda, int64_t n, double f, hipStream_t s);
// A(ri[e], ci[e]) = v[e] for e < z (0-based, distinct positions, all inside A)
int k_scatter(double* A, int64_t lda, const int64_t* ri, const int64_t* ci, const double* v, int64_t z,
              hipStream_t s);
// A <- A + A^T off the diagonal, A(i, i) <- (A(i, i) + A(i, i))/2 - sigma
int k_mm_symmetrize(double* A, int64_t lda, int64_t n, double sigma, hipStream_t s);
template <typename T>
int k_chol_leaf(T* A, int64_t lda, int64_t strideA, T* Li, int64_t ldl, int64_t strideL, int nb, int64_t col0,
                int* info, int batch, hipStream_t s);

// Structured Householder QR (qr.hip).
constexpr int kQrNb = 64;   // panel width
template <typename T>
// Panel at column p of width jb over rows [p, r1_end) and [n, r2_end) of each 2n x n (or n x n) matrix.
int k_qr_panel_step(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                    int64_t j, int apply_prev, int reduce_cur, double* part, int nwg, T* tau, int64_t strideTau,
                    int batch, hipStream_t s);
// Whole panel in one workgroup per shift: rows n + jb <= kQrPanelWgRows, width
// jb <= kQrPanelWgCols<T> (64 for FP32, 32 for FP64).
constexpr int kQrPanelWgRows = 576;
template <typename T>
constexpr int kQrPanelWgCols = sizeof(T) == 4 ? 64 : 32;
template <typename T>
int k_qr_panel_wg(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, T* tau, int64_t strideTau,
                  int batch, hipStream_t s);
// The whole panel in one cooperative launch per group of shifts (one grid-wide
// barrier per column); *done = false when the grid would not stay resident,
// and the caller then takes k_qr_panel_step.  part: >= 2 * batch * (nwg * 64 + 64)
// doubles with nwg = ceil(rows / 512).
// Shift-per-XCD QR panels with device-scope exchange stores even where the
// L2-local ones apply (1), or the automatic choice (0).
void qr_set_panel_device_stores(int on);
int qr_panel_device_stores();
template <typename T>
int k_qr_panel_coop(T* S, int64_t lds, int64_t strideS, int64_t n, int64_t p, int jb, int64_t r1_end, int64_t r2_end,
                    double* part, int64_t part_doubles, T* tau, int64_t strideTau, int batch, hipStream_t s,
                    bool* done);
template <typename T>
int k_qr_extract_v(const T* S, int64_t lds, int64_t strideS, T* V, int64_t ldv, int64_t strideV, int64_t p, int jb,
                   int64_t n, int64_t r1_end, int64_t r2_end, int batch, hipStream_t s);
template <typename T>
int k_qr_form_t(const T* G, int64_t ldg, int64_t strideG, const T* tau, int64_t strideTau, int64_t p, int jb, T* Tm,
                int64_t ldt, int64_t strideT, int batch, hipStream_t s);
template <typename T>
int k_set_identity_top(T* Q, int64_t ldq, int64_t strideQ, int64_t m, int64_t n, int batch, hipStream_t s);

}  // namespace zolo
