// Singular-value bounds of the polar iterations on the GPU:
//  * alpha = ||A||_F                                   (test_zolodrv1.f90:179)
//  * LU with partial pivoting of A/alpha               (PDGETRF, :185)
//  * ||(A/alpha)^{-1}||_1 by Higham's estimator (dlacn2, the algorithm behind
//    PDGECON('1') at :186); each estimator step is a pair of blocked
//    triangular solves on the GPU, the sign/argmax logic runs on the host
//  * the 2-norm power iteration of PDNORMEST (pdnormest.f90:43-80).
//
// LU panel: each 64-column panel is factored in one launch with the rows in
// registers (lu_panel_coop_kernel; one barrier per column among its
// workgroups), then one launch applies the panel's row interchanges to the
// other columns.  Matrices too tall for one resident grid take the per-column
// kernels (a launch that swaps the pivot row and one that applies the rank-1
// update while reducing the next column's argmax).  All reductions use
// per-workgroup partials combined in a fixed order.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "driver.hpp"
#include "grid_sync.hpp"

namespace zolo {
namespace {

constexpr int kT = 256;
constexpr int kLuNb = 64;
constexpr int kTb = 64;  // triangular-solve block

struct Piv {
  double v;
  int64_t i;
};

__device__ inline bool piv_better(double v, int64_t i, double bv, int64_t bi) {
  return v > bv || (v == bv && i < bi);
}

// Block-wide argmax of (|value|, row) pairs, ties to the smaller row (as IDAMAX).
__device__ inline void argmax_reduce(double bv, int64_t bi, Piv* out) {
  __shared__ double sv[kT];
  __shared__ int64_t si[kT];
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && piv_better(sv[threadIdx.x + s], si[threadIdx.x + s], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + s];
      si[threadIdx.x] = si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out->v = sv[0]; out->i = si[0]; }
}

// One thread per row: argmax of |A[i][j]| over rows i >= j.
__global__ __launch_bounds__(kT) void lu_argmax_kernel(const double* A, int64_t lda, int64_t n, int64_t j, int nwg,
                                                       Piv* part) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  double bv = -1.0;
  int64_t bi = INT64_MAX;
  if (i >= j && i < n) { bv = fabs(A[i + j * lda]); bi = i; }
  argmax_reduce(bv, bi, &part[blockIdx.x]);
}

// Reduce the partials, swap rows j <-> piv over all n columns, record ipiv.
__global__ __launch_bounds__(kT) void lu_swap_kernel(double* A, int64_t lda, int64_t n, int64_t j, int nwg,
                                                     const Piv* part, int64_t* ipiv, int* info) {
  __shared__ int64_t piv;
  __shared__ double sv[kT];
  __shared__ int64_t si[kT];
  {  // reduce the nwg partials with the whole block (fixed order)
    double bv = -1.0;
    int64_t bi = INT64_MAX;
    for (int k = threadIdx.x; k < nwg; k += kT)
      if (piv_better(part[k].v, part[k].i, bv, bi)) { bv = part[k].v; bi = part[k].i; }
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s && piv_better(sv[threadIdx.x + s], si[threadIdx.x + s], sv[threadIdx.x], si[threadIdx.x])) {
        sv[threadIdx.x] = sv[threadIdx.x + s];
        si[threadIdx.x] = si[threadIdx.x + s];
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    const double bv = sv[0];
    int64_t bi = si[0];
    if (bi == INT64_MAX) bi = j;
    piv = bi;
    if (blockIdx.x == 0) {
      ipiv[j] = bi;
      if (bv == 0.0 && info && *info == 0) *info = (int)(j + 1);
    }
  }
  __syncthreads();
  const int64_t pv = piv;
  if (pv == j) return;
  for (int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x; c < n; c += (int64_t)gridDim.x * kT) {
    const double a = A[j + c * lda];
    A[j + c * lda] = A[pv + c * lda];
    A[pv + c * lda] = a;
  }
}

// One thread per row i > j: l = A[i][j]/A[j][j]; A[i][c] -= l*A[j][c] for c in
// (j, pe), with the row held in registers; then the argmax partial of column
// j+1 from the updated values.
__global__ __launch_bounds__(kT) void lu_update_kernel(double* A, int64_t lda, int64_t n, int64_t j, int64_t pe,
                                                       int nwg, Piv* part) {
  __shared__ double prow[kLuNb];
  const int nc = (int)(pe - j - 1);
  if ((int)threadIdx.x < nc) prow[threadIdx.x] = A[j + (j + 1 + threadIdx.x) * lda];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  const double d = A[j + j * lda];
  const bool live = i > j && i < n;
  double v[kLuNb];
#pragma unroll
  for (int c = 0; c < kLuNb; ++c) v[c] = (live && c < nc) ? A[i + (j + 1 + c) * lda] : 0.0;
  if (live && d != 0.0) {
    const double l = A[i + j * lda] / d;
    A[i + j * lda] = l;
#pragma unroll
    for (int c = 0; c < kLuNb; ++c)
      if (c < nc) v[c] -= l * prow[c];
#pragma unroll
    for (int c = 0; c < kLuNb; ++c)
      if (c < nc) A[i + (j + 1 + c) * lda] = v[c];
  }
  if (j + 1 < pe) {
    double bv = -1.0;
    int64_t bi = INT64_MAX;
    if (live) { bv = fabs(v[0]); bi = i; }
    argmax_reduce(bv, bi, &part[blockIdx.x]);
  }
}

// Whole LU panel (columns [p, p + jb), jb <= 64, rows [p, n)) in one launch:
// one thread per row keeps the row's panel values in registers, shifted so
// that the current column is always v[0], and stores each column when it is
// final.  Per column: the workgroup's argmax of |v[0]| over rows >= j (ties to
// the smaller row, as IDAMAX) and that row go to the workgroup's slot, row j
// to a shared slot; one barrier among the workgroups; every workgroup picks
// the winner of the slots in the same order, swaps the pivot row and row j in
// registers and applies the rank-1 update (dgetf2).  Columns that are already
// final, and the columns outside the panel, get the row interchanges
// afterwards (lu_laswp_panel_kernel).  Slots alternate between two buffers;
// bar is a zeroed counter, or null when nwg == 1.
constexpr int kLuCoopT = 512;
constexpr int kLuSlot = kLuNb + 2;  // candidate row, |value|, row index
constexpr int kLuMaxWg = 64;        // workgroups of one panel: the winner is reduced by one wave
// doubles of the panel slots (two buffers) followed by the barrier counter
constexpr int64_t kLuSlotDoubles = 2 * ((int64_t)kLuMaxWg * kLuSlot + kLuNb) + 8;

__global__ __launch_bounds__(kLuCoopT) void lu_panel_coop_kernel(double* A, int64_t lda, int64_t n, int64_t p, int jb,
                                                                 double* slots, unsigned* bar, int nwg,
                                                                 int64_t* ipiv, int* info) {
  __shared__ double wbv[kLuCoopT / 64], rj[kLuNb], cand[kLuNb], rowj[kLuNb];
  __shared__ double crow[kLuMaxWg][kLuNb];  // every workgroup's candidate row
  __shared__ int64_t wbi[kLuCoopT / 64];
  __shared__ int64_t s_piv;
  __shared__ int s_slot;
  const int t = threadIdx.x, w = blockIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t i = p + (int64_t)w * kLuCoopT + t;
  const bool valid = i < n;
  const int64_t rec = (int64_t)nwg * kLuSlot + kLuNb;
  double v[kLuNb];
#pragma unroll
  for (int c = 0; c < kLuNb; ++c) v[c] = (valid && c < jb) ? A[i + (p + c) * lda] : 0.0;
  double* out = A + i + p * lda;
  for (int jj = 0; jj < jb; ++jj) {
    const int64_t j = p + jj;
    const int rem = jb - jj;
    double* G = slots + (int64_t)(jj & 1) * rec;
    double bv = -1.0;
    int64_t bi = INT64_MAX;
    if (valid && i >= j) { bv = fabs(v[0]); bi = i; }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int64_t oi = __shfl_xor(bi, o, 64);
      if (piv_better(ov, oi, bv, bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { wbv[wave] = bv; wbi[wave] = bi; }
    if (valid && i == j) {
#pragma unroll
      for (int c = 0; c < kLuNb; ++c) rowj[c] = v[c];
    }
    __syncthreads();
    // every thread reduces the waves' candidates in the same order (no extra barrier)
    double wb = wbv[0];
    int64_t wi = wbi[0];
    for (int k = 1; k < kLuCoopT / 64; ++k)
      if (piv_better(wbv[k], wbi[k], wb, wi)) { wb = wbv[k]; wi = wbi[k]; }
    if (valid && i == wi) {
#pragma unroll
      for (int c = 0; c < kLuNb; ++c) cand[c] = v[c];
    }
    __syncthreads();
    // wave 0 publishes the slots (the barrier's release is issued by its thread 0)
    if (t < kLuNb) {
      const bool has = wi != INT64_MAX;  // this workgroup has rows >= j
      if (t < rem) G[(int64_t)w * kLuSlot + t] = has ? cand[t] : 0.0;
      if (t < rem && (j - p) / kLuCoopT == w) G[(int64_t)nwg * kLuSlot + t] = rowj[t];
      if (t == 0) {
        G[(int64_t)w * kLuSlot + kLuNb] = wb;
        G[(int64_t)w * kLuSlot + kLuNb + 1] = (double)wi;  // row indices are exact in a double
      }
    }
    if (bar) grid_barrier(bar, (unsigned)(jj + 1) * (unsigned)nwg);
    else __syncthreads();
    // one round of loads: every candidate row and row j into LDS, while wave 0
    // reduces the candidates' (value, row) pairs
    for (int e = t; e < nwg * kLuNb; e += kLuCoopT) {
      const int k = e / kLuNb, c = e % kLuNb;
      crow[k][c] = (c < rem) ? G[(int64_t)k * kLuSlot + c] : 0.0;
    }
    if (t < kLuNb) rj[t] = (t < rem) ? G[(int64_t)nwg * kLuSlot + t] : 0.0;
    if (t < 64) {  // winner among the workgroups' candidates (nwg <= 64), fixed order
      double b = -1.0;
      int64_t b_i = INT64_MAX;
      int slot = 0;
      if (t < nwg) {
        b = G[(int64_t)t * kLuSlot + kLuNb];
        b_i = (int64_t)G[(int64_t)t * kLuSlot + kLuNb + 1];
        slot = t;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(b, o, 64);
        const int64_t oi = __shfl_xor(b_i, o, 64);
        const int os = __shfl_xor(slot, o, 64);
        if (piv_better(ob, oi, b, b_i)) { b = ob; b_i = oi; slot = os; }
      }
      if (t == 0) {
        s_piv = b_i;
        s_slot = slot;
        if (w == 0) {
          ipiv[j] = b_i;
          if (b == 0.0 && *info == 0) *info = (int)(j + 1);
        }
      }
    }
    __syncthreads();
    const int64_t piv = s_piv;
    const int sl = s_slot;  // crow[sl] is the pivot row
    if (valid && piv != j) {
      if (i == j) {
#pragma unroll
        for (int c = 0; c < kLuNb; ++c) v[c] = crow[sl][c];
      } else if (i == piv) {
#pragma unroll
        for (int c = 0; c < kLuNb; ++c) v[c] = rj[c];
      }
    }
    const double d = crow[sl][0];
    if (valid && i > j && d != 0.0) {
      const double l = v[0] / d;
      v[0] = l;
#pragma unroll
      for (int c = 1; c < kLuNb; ++c) v[c] -= l * crow[sl][c];
    }
    if (valid) *out = v[0];
    out += lda;
#pragma unroll
    for (int c = 0; c + 1 < kLuNb; ++c) v[c] = v[c + 1];
    v[kLuNb - 1] = 0.0;
    // No barrier here: the next column writes wbv/wbi, rowj and cand only
    // after this column's last barrier (every reader has passed it), and crow,
    // rj, s_piv, s_slot only after the next grid barrier.
  }
}

// Row interchanges of one factored panel (pivots ipiv[p .. p + jb), in order)
// for every column c of the matrix except what the panel kernel already
// swapped: panel column p + cc only takes the interchanges of steps after cc.
// One thread per column.
__global__ __launch_bounds__(kT) void lu_laswp_panel_kernel(double* A, int64_t lda, int64_t n, int64_t p, int jb,
                                                            const int64_t* ipiv) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= n) return;
  const int64_t first = (c >= p && c < p + jb) ? c - p + 1 : 0;
  double* col = A + c * lda;
  for (int64_t jj = first; jj < jb; ++jj) {
    const int64_t j = p + jj, pv = ipiv[j];
    if (pv != j) {
      const double a = col[j];
      col[j] = col[pv];
      col[pv] = a;
    }
  }
}

// Unit-lower inverse of the jb x jb diagonal block (one workgroup): thread c
// forms column c by forward substitution, the block and the result in LDS.
__global__ __launch_bounds__(kT) void unit_lower_inv_kernel(const double* A, int64_t lda, int jb, double* Li,
                                                            int64_t ldl) {
  __shared__ double a[kLuNb][kLuNb + 1];
  __shared__ double x[kLuNb][kLuNb + 1];
  for (int e = threadIdx.x; e < jb * jb; e += kT) {
    const int i = e % jb, k = e / jb;
    a[i][k] = A[i + k * lda];
    x[i][k] = (i == k) ? 1.0 : 0.0;
  }
  __syncthreads();
  // forward substitution for all columns at once (axpy form): at step k, row k
  // of X is final and updates the rows below; each element subtracts the same
  // terms in the same order as a per-column substitution
  for (int k = 0; k + 1 < jb; ++k) {
    const int m = jb - k - 1;
    for (int e = threadIdx.x; e < m * (k + 1); e += kT) {
      const int i = k + 1 + e % m, c = e / m;
      x[i][c] -= a[i][k] * x[k][c];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < jb * jb; e += kT) {
    const int i = e % jb, k = e / jb;
    Li[i + k * ldl] = x[i][k];
  }
}

__global__ __launch_bounds__(kT) void copy2d_kernel(const double* S, int64_t lds, double* D, int64_t ldd, int64_t m,
                                                    int64_t n) {
  const int64_t total = m * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m, j = e / m;
    D[i + j * ldd] = S[i + j * lds];
  }
}

// ---- triangular solves with one right-hand side --------------------------
// op(T)[i][k] = trans ? M[k + i*ld] : M[i + k*ld]

// Inverses of the kTb x kTb diagonal blocks of the LU factors, one workgroup
// per block: Linv_b = L_bb^{-1} (unit lower), Uinv_b = U_bb^{-1} (upper).
// Thread c forms column c by substitution in LDS.
__global__ __launch_bounds__(kTb) void lu_block_inv_kernel(const double* LU, int64_t ld, int64_t n, double* Linv,
                                                           double* Uinv) {
  __shared__ double a[kTb][kTb + 1];
  __shared__ double xl[kTb][kTb + 1];
  __shared__ double xu[kTb][kTb + 1];
  const int64_t b0 = (int64_t)blockIdx.x * kTb;
  const int bs = (int)min((int64_t)kTb, n - b0);
  const int t = threadIdx.x;
  for (int e = t; e < kTb * kTb; e += kTb) {
    const int i = e % kTb, k = e / kTb;
    a[i][k] = (i < bs && k < bs) ? LU[(b0 + i) + (b0 + k) * ld] : (i == k ? 1.0 : 0.0);
  }
  __syncthreads();
  const int c = t;
  // L x = e_c (unit lower): x[i] = -sum_{k in [c, i)} a[i][k] x[k], x[c] = 1
  for (int i = 0; i < kTb; ++i) xl[i][c] = (i == c) ? 1.0 : 0.0;
  for (int i = c + 1; i < kTb; ++i) {
    double sum = 0.0;
    for (int k = c; k < i; ++k) sum += a[i][k] * xl[k][c];
    xl[i][c] = -sum;
  }
  // U x = e_c (upper): x[i] = (delta_ic - sum_{k in (i, c]} a[i][k] x[k]) / a[i][i]
  for (int i = 0; i < kTb; ++i) xu[i][c] = 0.0;
  for (int i = c; i >= 0; --i) {
    double sum = (i == c) ? 1.0 : 0.0;
    for (int k = i + 1; k <= c; ++k) sum -= a[i][k] * xu[k][c];
    xu[i][c] = sum / a[i][i];
  }
  __syncthreads();
  double* Lb = Linv + (int64_t)blockIdx.x * kTb * kTb;
  double* Ub = Uinv + (int64_t)blockIdx.x * kTb * kTb;
  for (int e = t; e < kTb * kTb; e += kTb) {
    const int i = e % kTb, k = e / kTb;
    Lb[i + k * kTb] = xl[i][k];
    Ub[i + k * kTb] = xu[i][k];
  }
}

// One block of a blocked triangular solve: every workgroup forms
// x_b' = op(D_b)^{-1} x_b from the block's explicit inverse (Dinv, op =
// transposed when dtrans), workgroup 0 writes it to y, and rows r in [r0, r1)
// get x[r] -= sum_k op(M)[r][b0+k] x_b'[k].  x_b itself is not written, so the
// workgroups never race on it.  op(M) = M (column-oriented update): 64 rows per
// pass, the 64 block columns split in four quarters summed through LDS.
// op(M) = M^T: 32 rows per pass, eight lanes per row reading the row's 64
// contiguous values of M (k = lane + 8 e) and summed by shuffles.
constexpr int kTrsvRowsN = 64, kTrsvRowsT = 32;
__global__ __launch_bounds__(kT) void trsv_step_kernel(const double* M, int64_t ld, int trans, const double* Dinv,
                                                       int dtrans, double* x, double* y, int64_t b0, int bs,
                                                       int64_t r0, int64_t r1) {
  __shared__ double xb[kTb], xs[kTb];
  __shared__ double part[kT / kTrsvRowsN][kTrsvRowsN];
  const int t = threadIdx.x;
  if (t < bs) xb[t] = x[b0 + t];
  __syncthreads();
  if (t < kTb) {
    double sum = 0.0;
    if (t < bs)
      for (int k = 0; k < bs; ++k) sum += (dtrans ? Dinv[k + t * kTb] : Dinv[t + k * kTb]) * xb[k];
    xs[t] = sum;  // zero past the block
    if (blockIdx.x == 0 && t < bs) y[b0 + t] = sum;
  }
  __syncthreads();
  if (!trans) {
    const int rl = t % kTrsvRowsN, q = t / kTrsvRowsN;
    for (int64_t base = r0 + (int64_t)blockIdx.x * kTrsvRowsN; base < r1; base += (int64_t)gridDim.x * kTrsvRowsN) {
      const int64_t r = base + rl;
      double sum = 0.0;
      if (r < r1) {
        const double* Mr = M + r + (b0 + q * 16) * ld;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
          if (q * 16 + kk < bs) sum += Mr[kk * ld] * xs[q * 16 + kk];
      }
      part[q][rl] = sum;
      __syncthreads();
      if (q == 0 && r < r1) x[r] -= (part[0][rl] + part[1][rl]) + (part[2][rl] + part[3][rl]);
      __syncthreads();
    }
  } else {
    const int rl = t >> 3, q = t & 7;
    for (int64_t base = r0 + (int64_t)blockIdx.x * kTrsvRowsT; base < r1; base += (int64_t)gridDim.x * kTrsvRowsT) {
      const int64_t r = base + rl;
      double sum = 0.0;
      if (r < r1) {
        const double* Mc = M + b0 + r * ld;
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (q + 8 * e < bs) sum += Mc[q + 8 * e] * xs[q + 8 * e];
      }
      sum += __shfl_xor(sum, 1, 8);
      sum += __shfl_xor(sum, 2, 8);
      sum += __shfl_xor(sum, 4, 8);
      if (q == 0 && r < r1) x[r] -= sum;
    }
  }
}

// ---- GEMV for the power iteration ------------------------------------------
// y = A x, column chunks of 256 -> partial[chunk][row]
__global__ __launch_bounds__(kT) void gemv_n_kernel(const double* A, int64_t lda, int64_t m, int64_t n,
                                                    const double* x, double* partial) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.y * kT, c1 = min(n, c0 + kT);
  if (r >= m) return;
  double s = 0.0;
  for (int64_t c = c0; c < c1; ++c) s += A[r + c * lda] * x[c];
  partial[blockIdx.y * m + r] = s;
}
__global__ __launch_bounds__(kT) void gemv_n_reduce_kernel(const double* partial, int64_t m, int nchunk, double* y) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= m) return;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += partial[k * m + r];
  y[r] = s;
}
// y = A^T x: one workgroup per column
__global__ __launch_bounds__(kT) void gemv_t_kernel(const double* A, int64_t lda, int64_t m, const double* x,
                                                    double* y) {
  __shared__ double sh[kT];
  const int64_t c = blockIdx.x;
  double s = 0.0;
  for (int64_t r = threadIdx.x; r < m; r += kT) s += A[r + c * lda] * x[r];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = kT / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[c] = sh[0];
}

inline unsigned grid_for(int64_t total) {
  int64_t g = (total + kT - 1) / kT;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

// One thread per row in the panel kernels.
int lu_nwg(int64_t n) { return (int)std::max<int64_t>(1, (n + kT - 1) / kT); }

// Triangular solve with the unit-lower kLuNb-wide diagonal block at (p, p):
// B (jb x nc, rows p.., ld lda) <- L11^{-1} B, through the block's explicit
// inverse and a GEMM into W, copied back.
int lu_trsm_block(const Ctx& c, double* A, int64_t lda, int64_t p, int jb, double* B, int64_t nc, double* L11i,
                  double* W) {
  hipLaunchKernelGGL(unit_lower_inv_kernel, dim3(1), dim3(kT), 0, c.s, A + p + p * lda, lda, jb, L11i, (int64_t)kLuNb);
  GemmDesc d;  // W = L11^{-1} B
  d.transA = 0; d.transB = 0; d.M = jb; d.N = nc; d.K = jb;
  d.A = L11i; d.lda = kLuNb; d.B = B; d.ldb = lda; d.C = W; d.ldc = jb; d.aHi = 0;
  ZOLO_CHECK(gemm(Prec::F64, d, c.s));
  hipLaunchKernelGGL(copy2d_kernel, dim3(grid_for(jb * nc)), dim3(kT), 0, c.s, W, (int64_t)jb, B, lda, (int64_t)jb, nc);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

// In-place LU with partial pivoting (row swaps applied to full rows), two
// levels: panels of kLuNb columns factored column by column and applied only
// inside their outer block of kLuOuter columns; each outer block then updates
// the trailing matrix with K = kLuOuter (A22 -= L21 U12 on MFMA).
constexpr int kLuOuter = 256;

// Panel [p, p + jb) in one launch of lu_panel_coop_kernel followed by the
// row interchanges of the other columns; *done = false when the rows do not
// fit one resident grid (the caller then takes the per-column kernels).
int lu_panel_coop(const Ctx& c, double* A, int64_t lda, int64_t n, int64_t p, int jb, int64_t* ipiv, int* info,
                  double* slots, bool* done) {
  *done = false;
  static int cap[64] = {0};
  int dev = 0;
  ZOLO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return kOk;
  if (cap[dev] == 0) {
    int per_cu = 0, cus = 0, coop = 0;
    ZOLO_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    ZOLO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    ZOLO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)lu_panel_coop_kernel, kLuCoopT, 0));
    cap[dev] = coop ? per_cu * cus : -1;
    if (cap[dev] <= 0) cap[dev] = -1;
  }
  const int64_t rows = n - p;
  int nwg = (int)((rows + kLuCoopT - 1) / kLuCoopT);
  if (nwg > kLuMaxWg) return kOk;
  unsigned* bar = reinterpret_cast<unsigned*>(slots + 2 * ((int64_t)kLuMaxWg * kLuSlot + kLuNb));
  if (nwg == 1) {
    unsigned* nobar = nullptr;
    hipLaunchKernelGGL(lu_panel_coop_kernel, dim3(1), dim3(kLuCoopT), 0, c.s, A, lda, n, p, jb, slots, nobar, nwg,
                       ipiv, info);
    ZOLO_HIP(hipGetLastError());
  } else {
    if (cap[dev] < nwg) return kOk;
    ZOLO_HIP(hipMemsetAsync(bar, 0, sizeof(unsigned), c.s));
    void* args[] = {&A, &lda, &n, &p, &jb, &slots, &bar, &nwg, &ipiv, &info};
    ZOLO_HIP(hipLaunchCooperativeKernel((const void*)lu_panel_coop_kernel, dim3(nwg), dim3(kLuCoopT), args, 0, c.s));
  }
  hipLaunchKernelGGL(lu_laswp_panel_kernel, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, c.s, A, lda, n, p, jb,
                     (const int64_t*)ipiv);
  ZOLO_HIP(hipGetLastError());
  *done = true;
  return kOk;
}

int getrf(const Ctx& c, double* A, int64_t lda, int64_t n, int64_t* ipiv, int* info, Piv* part, double* L11i,
          double* W, double* slots) {
  const int nwg = lu_nwg(n);
  for (int64_t P = 0; P < n; P += kLuOuter) {
    const int JB = (int)std::min<int64_t>(kLuOuter, n - P);
    const int64_t PE = P + JB;
    for (int64_t p = P; p < PE; p += kLuNb) {
      const int jb = (int)std::min<int64_t>(kLuNb, PE - p);
      const int64_t pe = p + jb;
      bool done = false;
      ZOLO_CHECK(lu_panel_coop(c, A, lda, n, p, jb, ipiv, info, slots, &done));
      if (!done) {
        hipLaunchKernelGGL(lu_argmax_kernel, dim3(nwg), dim3(kT), 0, c.s, A, lda, n, p, nwg, part);
        for (int64_t j = p; j < pe; ++j) {
          hipLaunchKernelGGL(lu_swap_kernel, dim3(grid_for(n)), dim3(kT), 0, c.s, A, lda, n, j, nwg, part, ipiv,
                             info);
          hipLaunchKernelGGL(lu_update_kernel, dim3(nwg), dim3(kT), 0, c.s, A, lda, n, j, pe, nwg, part);
        }
      }
      ZOLO_HIP(hipGetLastError());
      const int64_t nc = PE - pe;  // rest of the outer block
      if (nc <= 0) continue;
      ZOLO_CHECK(lu_trsm_block(c, A, lda, p, jb, A + p + pe * lda, nc, L11i, W));
      GemmDesc d;  // A[pe:, pe:PE] -= L21 U12
      d.transA = 0; d.transB = 0; d.M = n - pe; d.N = nc; d.K = jb;
      d.A = A + pe + p * lda; d.lda = lda; d.B = A + p + pe * lda; d.ldb = lda; d.C = A + pe + pe * lda; d.ldc = lda;
      d.alpha = -1.0; d.beta = 1.0;
      ZOLO_CHECK(gemm(Prec::F64, d, c.s));
    }
    const int64_t nc = n - PE;
    if (nc <= 0) continue;
    // U12 = L11^{-1} A12 by blocks of kLuNb rows (forward), then A22 -= L21 U12
    for (int64_t b = P; b < PE; b += kLuNb) {
      const int bs = (int)std::min<int64_t>(kLuNb, PE - b);
      ZOLO_CHECK(lu_trsm_block(c, A, lda, b, bs, A + b + PE * lda, nc, L11i, W));
      if (b + bs < PE) {
        GemmDesc d;  // A12[b+bs:PE] -= L[b+bs:PE, b:b+bs] U12[b:b+bs]
        d.transA = 0; d.transB = 0; d.M = PE - (b + bs); d.N = nc; d.K = bs;
        d.A = A + (b + bs) + b * lda; d.lda = lda; d.B = A + b + PE * lda; d.ldb = lda;
        d.C = A + (b + bs) + PE * lda; d.ldc = lda; d.alpha = -1.0; d.beta = 1.0;
        ZOLO_CHECK(gemm(Prec::F64, d, c.s));
      }
    }
    GemmDesc d;  // A22 -= L21 U12   (K = JB)
    d.transA = 0; d.transB = 0; d.M = nc; d.N = nc; d.K = JB;
    d.A = A + PE + P * lda; d.lda = lda; d.B = A + P + PE * lda; d.ldb = lda; d.C = A + PE + PE * lda; d.ldc = lda;
    d.alpha = -1.0; d.beta = 1.0;
    ZOLO_CHECK(gemm(Prec::F64, d, c.s));
  }
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

// y <- op(T)^{-1} x, blocked, with the explicit diagonal-block inverses of the
// LU factors (x is used as workspace).  op(T): L (lower, unit), U (upper),
// U^T (lower), L^T (upper, unit).
int trsv(const Ctx& c, const double* M, int64_t ld, int64_t n, int trans, int lower, const double* Linv,
         const double* Uinv, double* x, double* y) {
  const int64_t nb = (n + kTb - 1) / kTb;
  const bool useL = (lower && !trans) || (!lower && trans);  // op(T) built from L
  const double* Dinv = useL ? Linv : Uinv;
  for (int64_t s = 0; s < nb; ++s) {
    const int64_t b = lower ? s : nb - 1 - s;
    const int64_t b0 = b * kTb;
    const int bs = (int)std::min<int64_t>(kTb, n - b0);
    const int64_t r0 = lower ? b0 + bs : 0, r1 = lower ? n : b0;
    const int64_t per = trans ? kTrsvRowsT : kTrsvRowsN;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((r1 - r0 + per - 1) / per, 8192));
    hipLaunchKernelGGL(trsv_step_kernel, dim3(g), dim3(kT), 0, c.s, M, ld, trans, Dinv + b * kTb * kTb, trans, x, y,
                       b0, bs, r0, r1);
  }
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

// Higham's estimate of ||M^{-1}||_1 (LAPACK dlacn2 as xGECON / xTRCON drive
// it, with the final alternating-sign vector test); solve(h, trans) replaces
// the host vector h by M^{-1} h (trans = 0) or M^{-T} h (trans = 1).
template <typename Solve>
int lacn2_inv(int64_t n, Solve&& solve, double* est_out) {
  std::vector<double> h(n), sgn(n);
  auto asum = [&]() { double t = 0; for (int64_t i = 0; i < n; ++i) t += std::fabs(h[i]); return t; };
  auto iamax = [&]() { int64_t b = 0; for (int64_t i = 1; i < n; ++i) if (std::fabs(h[i]) > std::fabs(h[b])) b = i; return b; };
  double est;
  std::fill(h.begin(), h.end(), 1.0 / n);
  ZOLO_CHECK(solve(h, 0));
  if (n == 1) {
    est = std::fabs(h[0]);
  } else {
    est = asum();
    for (int64_t i = 0; i < n; ++i) sgn[i] = h[i] >= 0 ? 1.0 : -1.0;
    h = sgn;
    ZOLO_CHECK(solve(h, 1));
    int64_t j = iamax();
    int iter = 2;
    while (true) {
      std::fill(h.begin(), h.end(), 0.0);
      h[j] = 1.0;
      ZOLO_CHECK(solve(h, 0));
      const double estold = est;
      est = asum();
      bool same = true;
      for (int64_t i = 0; i < n; ++i)
        if ((h[i] >= 0 ? 1.0 : -1.0) != sgn[i]) { same = false; break; }
      if (same || est <= estold) break;
      for (int64_t i = 0; i < n; ++i) sgn[i] = h[i] >= 0 ? 1.0 : -1.0;
      h = sgn;
      ZOLO_CHECK(solve(h, 1));
      const int64_t jlast = j;
      j = iamax();
      if (h[jlast] != std::fabs(h[j]) && iter < 5) {
        ++iter;
        continue;
      }
      break;
    }
    for (int64_t i = 0; i < n; ++i) h[i] = ((i % 2 == 0) ? 1.0 : -1.0) * (1.0 + double(i) / double(n - 1));
    ZOLO_CHECK(solve(h, 0));
    const double temp = 2.0 * asum() / (3.0 * n);
    if (temp > est) est = temp;
  }
  *est_out = est;
  return kOk;
}

}  // namespace

size_t bounds_bytes(int64_t n) {
  const int64_t nb = (n + kTb - 1) / kTb;
  return sizeof(double) * (size_t)(n * n + 3 * n + kLuNb * kLuNb + kLuNb * n + 8 + 2 * nb * kTb * kTb +
                                  kLuSlotDoubles) +
         sizeof(int64_t) * n + sizeof(Piv) * (size_t)(n / kT + 2) + 17 * 256;
}

int bounds_run(int convention, int64_t n, const double* A, int64_t lda, double* alpha, double* beta, Arena& ar,
               hipStream_t s) {
  Ctx c;
  c.s = s;
  double* LU = ar.take<double>((size_t)n * n);
  double* x = ar.take<double>(n);
  double* y = ar.take<double>(n);
  double* col = ar.take<double>(n);
  const int64_t nblk = (n + kTb - 1) / kTb;
  double* Linv = ar.take<double>((size_t)nblk * kTb * kTb);
  double* Uinv = ar.take<double>((size_t)nblk * kTb * kTb);
  double* L11i = ar.take<double>(kLuNb * kLuNb);
  double* W = ar.take<double>((size_t)kLuNb * n);
  double* out = ar.take<double>(8);
  int64_t* ipiv = ar.take<int64_t>(n);
  Piv* part = ar.take<Piv>((size_t)(n / kT + 2));
  int* dinfo = ar.take<int>(4);
  double* slots = ar.take<double>((size_t)kLuSlotDoubles);
  ZOLO_ARENA_OK(ar);
  // alpha = ||A||_F ; X = A/alpha
  ZOLO_CHECK(k_sumsq<double>(A, lda, n, n, col, (int)std::min<int64_t>(n, 1024), out, s));
  double a2;
  ZOLO_HIP(hipMemcpyAsync(&a2, out, sizeof(double), hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  const double al = std::sqrt(a2);
  if (!(al > 0.0)) {
    set_error("bounds: zero matrix");
    return kErrArg;
  }
  ZOLO_CHECK(k_scale_copy<double, double>(A, lda, LU, n, n, n, 1.0 / al, s));
  double xnorm1 = 1.0;
  if (convention == 1) {
    ZOLO_CHECK(k_norm1<double>(LU, n, n, n, col, out, s));
    ZOLO_HIP(hipMemcpyAsync(&xnorm1, out, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  ZOLO_HIP(hipMemsetAsync(dinfo, 0, sizeof(int), s));
  ZOLO_CHECK(getrf(c, LU, n, n, ipiv, dinfo, part, L11i, W, slots));
  int hinfo = 0;
  ZOLO_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  if (hinfo != 0) {
    set_error("bounds: matrix is exactly singular (LU pivot " + std::to_string(hinfo) + ")");
    return kErrArg;
  }
  hipLaunchKernelGGL(lu_block_inv_kernel, dim3((unsigned)nblk), dim3(kTb), 0, s, LU, n, n, Linv, Uinv);
  ZOLO_HIP(hipGetLastError());
  // dlacn2 on inv(U) inv(L) (xGECON applies no permutation).
  auto solve = [&](std::vector<double>& h, int trans) -> int {
    ZOLO_HIP(hipMemcpyAsync(x, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (!trans) {
      ZOLO_CHECK(trsv(c, LU, n, n, 0, 1, Linv, Uinv, x, y));   // L (unit lower)
      ZOLO_CHECK(trsv(c, LU, n, n, 0, 0, Linv, Uinv, y, x));   // U
    } else {
      ZOLO_CHECK(trsv(c, LU, n, n, 1, 1, Linv, Uinv, x, y));   // U^T (lower)
      ZOLO_CHECK(trsv(c, LU, n, n, 1, 0, Linv, Uinv, y, x));   // L^T (unit upper)
    }
    ZOLO_HIP(hipMemcpyAsync(h.data(), x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    ZOLO_HIP(hipStreamSynchronize(s));
    return kOk;
  };
  double est = 0.0;
  ZOLO_CHECK(lacn2_inv(n, solve, &est));
  const double rcond = 1.0 / est;
  *alpha = al;
  *beta = xnorm1 * rcond / std::sqrt((double)n);
  return kOk;
}

size_t rbound_bytes(int64_t n) {
  const int64_t nb = (n + kTb - 1) / kTb;
  return sizeof(double) * (size_t)(3 * n + 2 * nb * kTb * kTb) + 8 * 256;
}

int rbound_run(int64_t n, const double* R, int64_t ldr, double* est, Arena& ar, hipStream_t s) {
  Ctx c;
  c.s = s;
  const int64_t nblk = (n + kTb - 1) / kTb;
  double* x = ar.take<double>(n);
  double* y = ar.take<double>(n);
  double* Linv = ar.take<double>((size_t)nblk * kTb * kTb);  // formed from the reflectors, unused
  double* Uinv = ar.take<double>((size_t)nblk * kTb * kTb);
  ZOLO_ARENA_OK(ar);
  hipLaunchKernelGGL(lu_block_inv_kernel, dim3((unsigned)nblk), dim3(kTb), 0, s, R, ldr, n, Linv, Uinv);
  ZOLO_HIP(hipGetLastError());
  // dlacn2 on inv(R) (as DTRCON('1', 'U', 'N') drives it)
  auto solve = [&](std::vector<double>& h, int trans) -> int {
    ZOLO_HIP(hipMemcpyAsync(x, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    ZOLO_CHECK(trsv(c, R, ldr, n, trans, trans ? 1 : 0, Linv, Uinv, x, y));  // R, or R^T (lower)
    ZOLO_HIP(hipMemcpyAsync(h.data(), y, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    ZOLO_HIP(hipStreamSynchronize(s));
    return kOk;
  };
  return lacn2_inv(n, solve, est);
}

size_t normest_bytes(int64_t m, int64_t n) {
  const int64_t nch = (n + kT - 1) / kT;
  return sizeof(double) * (size_t)(m + n + nch * m + 1024 + 16) + 8 * 256;
}

int normest_run(int64_t m, int64_t n, const double* A, int64_t lda, double tol, double* est, int* cnt, Arena& ar,
                hipStream_t s) {
  double* x = ar.take<double>(n);
  double* sx = ar.take<double>(m);
  const int64_t nch = (n + kT - 1) / kT;
  double* part = ar.take<double>((size_t)nch * m);
  double* red = ar.take<double>(1024);
  double* out = ar.take<double>(16);
  ZOLO_ARENA_OK(ar);
  auto nrm = [&](const double* v, int64_t len, double* r) -> int {
    ZOLO_CHECK(k_sumsq<double>(v, len, len, 1, red, 1024, out, s));
    ZOLO_HIP(hipMemcpyAsync(r, out, sizeof(double), hipMemcpyDeviceToHost, s));
    ZOLO_HIP(hipStreamSynchronize(s));
    *r = std::sqrt(*r);
    return kOk;
  };
  ZOLO_CHECK(k_norm1<double>(A, lda, m, n, x, out, s));  // x = column abs sums
  double e;
  ZOLO_CHECK(nrm(x, n, &e));
  *cnt = 0;
  if (e == 0.0) {
    *est = 0.0;
    return kOk;
  }
  ZOLO_CHECK(k_scale_copy<double, double>(x, n, x, n, n, 1, 1.0 / e, s));
  double e0 = 0.0;
  const int maxiter = 100;
  int it = 0;
  while (std::fabs(e - e0) > tol * e && it < maxiter) {
    e0 = e;
    hipLaunchKernelGGL(gemv_n_kernel, dim3((unsigned)((m + kT - 1) / kT), (unsigned)nch), dim3(kT), 0, s, A, lda, m, n, x,
                       part);
    hipLaunchKernelGGL(gemv_n_reduce_kernel, dim3((unsigned)((m + kT - 1) / kT)), dim3(kT), 0, s, part, m, (int)nch, sx);
    ZOLO_HIP(hipGetLastError());
    double nsx, nx;
    ZOLO_CHECK(nrm(sx, m, &nsx));
    hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)n), dim3(kT), 0, s, A, lda, m, sx, x);
    ZOLO_HIP(hipGetLastError());
    ZOLO_CHECK(nrm(x, n, &nx));
    e = nx / nsx;
    ZOLO_CHECK(k_scale_copy<double, double>(x, n, x, n, n, 1, 1.0 / nx, s));
    ++it;
  }
  *est = e;
  *cnt = it;
  return kOk;
}

// ---------------------------------------------------------------------------
// PDNORMEST without allocation: the same power iteration with its scratch
// (x, S x, the mat-vec partials and the norm partials) in a fixed device
// array of this library instead of the workspace.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocBytes = size_t(16) << 20;  // 16 MiB per device
__device__ __align__(256) char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
char* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU
}  // namespace

}  // namespace zolo

// No counterpart in the reference: zolo_normest_device (PDNORMEST,
// pdnormest.f90:2) for callers that cannot allocate device memory.
extern "C" int zolo_normest_noalloc_device(int64_t m, int64_t n, const double* A, int64_t lda, double tol,
                                           double* est, int* cnt, void* stream) {
  using namespace zolo;
  if (m < 1) return -1;
  if (n < 1) return -2;
  if (!A) return -3;
  if (lda < m) return -4;
  if (!est) return -6;
  if (!cnt) return -7;
  if (normest_bytes(m, n) > kNoAllocBytes) {
    set_error("zolo_normest_noalloc_device: m = " + std::to_string(m) + ", n = " + std::to_string(n) + " needs " +
              std::to_string(normest_bytes(m, n)) + " bytes of scratch, more than the fixed " +
              std::to_string(kNoAllocBytes));
    return -1;
  }
  int dev = 0;
  ZOLO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("zolo_normest_noalloc_device: device index out of range");
    return kErrArg;
  }
  if (!g_scratch_ptr[dev]) {
    // the array's own name is a host-side stand-in; kernels get this address
    void* p = nullptr;
    ZOLO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  Arena ar(g_scratch_ptr[dev], kNoAllocBytes);
  return normest_run(m, n, A, lda, tol, est, cnt, ar, (hipStream_t)stream);
}

