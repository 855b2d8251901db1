// Kernels of the batched small-matrix polar decomposition: many independent
// square FP32 matrices of uneven size n_i <= kMaxBatchN, grouped into size
// buckets nb (n_i rounded up to a multiple of 32) and padded in HBM to
// diag(A_i, I).  Every kernel takes its matrix (or (matrix, shift) item)
// from blockIdx and the per-bucket tables built by the host.
//
//  * alpha_kernel        ||A_i||_F in FP64, one workgroup per matrix
//                        (test_zolodrv1.f90:179).
//  * lu_pad_kernel       W_i = diag(A_i / alpha_i, I) in FP64.
//  * lu_panel_wg_kernel  one 32-column panel of the LU with partial pivoting
//                        (PDGETRF at test_zolodrv1.f90:185), one workgroup per
//                        matrix with one row per thread in registers; it also
//                        forms the inverse of the unit-lower diagonal block and
//                        applies the panel's row interchanges to the other
//                        columns.  The trailing update is an MFMA GEMM.
//  * condest_kernel      Higham's 1-norm estimate of ||(A_i/alpha_i)^{-1}||_1
//                        from the LU (dlacn2 behind PDGECON at :186), one
//                        workgroup per matrix; beta_i = rcond / sqrt(n_i) (:189).
//  * pad_scale_kernel    X_i = diag(A_i / (alpha_i beta_i), I) in FP32 (or FP64).
//  * build_stack_items / shift_copy_items: the per-item inputs of the QR and
//                        Cholesky steps, from X or X^T X of the item's matrix.
//  * formx2_items_kernel X_i <- (X_i + sum of its items' Z) / FONE_i and
//                        ERR_i^2 (pzolopd1.f90:477-481, FormX2); formx2_f64_kernel
//                        the same for an FP64 step, rounding the result to FP32.
//  * gram_err_kernel     ||X_i^T X_i - I||_F^2 over the unpadded block.
//  * unpad_kernel        U_i = leading n_i x n_i block of X_i.
#include <algorithm>

#include "batched.hpp"

namespace zolo {
namespace {

constexpr int kT = 256;

__device__ inline double block_sum256(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

inline unsigned gfor(int64_t total, int64_t cap) {
  int64_t g = (total + kT - 1) / kT;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

__global__ __launch_bounds__(kT) void alpha_kernel(const BatchMat* mats, double* alpha) {
  __shared__ double sh[4];
  const BatchMat m = mats[blockIdx.x];
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < m.n * m.n; e += kT) {
    const double v = (double)m.A[(e % m.n) + (e / m.n) * m.lda];
    s += v * v;
  }
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) alpha[blockIdx.x] = sqrt(s);
}

// ---------------------------------------------------------------------------
// LU of the padded FP64 copies (bucket layout: matrix b at W + b * nb * nb).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void lu_pad_kernel(const BatchMat* mats, const int* idx, const double* alpha,
                                                    double* W, int64_t nb) {
  const int b = blockIdx.y;
  const int g = idx[b];
  const BatchMat m = mats[g];
  double* Wb = W + (int64_t)b * nb * nb;
  const double ia = 1.0 / alpha[g];
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < nb * nb; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % nb, j = e / nb;
    double v;
    if (i < m.n && j < m.n) v = ia * (double)m.A[i + j * m.lda];
    else v = (i == j) ? 1.0 : 0.0;
    Wb[e] = v;
  }
}

constexpr int kLuRows = kMaxBatchN;
constexpr int kLuNb = kBatchLuNb;

// Each thread keeps the not yet eliminated part of its row in registers,
// shifted so that the current column is always v[0]: no register is indexed
// by the loop counter.  Every step stores column j of all rows and shifts.
// Row interchanges exchange the register rows of the two threads at once; on
// the already stored panel columns and on the columns outside the panel they
// are applied after the panel, in order.
__global__ __launch_bounds__(kLuRows) void lu_panel_wg_kernel(double* W, int64_t nb, int64_t k0, double* Linv,
                                                              int* info) {
  __shared__ double wv[kLuRows / 64];
  __shared__ int wi[kLuRows / 64];
  __shared__ double urow[kLuNb], xrow[kLuNb];
  __shared__ int piv[kLuNb];
  __shared__ double L[kLuNb][kLuNb + 1];
  __shared__ double Xi[kLuNb][kLuNb + 1];
  const int b = blockIdx.x;
  double* A = W + (int64_t)b * nb * nb;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nwave = blockDim.x >> 6;
  const int64_t rows = nb - k0;
  const int jb = (int)(rows < kLuNb ? rows : kLuNb);
  const int64_t i = k0 + t;
  const bool mine = t < rows;
  double v[kLuNb];
#pragma unroll
  for (int c = 0; c < kLuNb; ++c) v[c] = (mine && c < jb) ? A[i + (k0 + c) * nb] : 0.0;
  for (int jj = 0; jj < jb; ++jj) {
    const int64_t j = k0 + jj;
    // argmax of |v[0]| over rows >= j, ties to the smaller row (IDAMAX)
    double bv = (mine && i >= j) ? fabs(v[0]) : -1.0;
    int bi = (int)i;
    for (int o = 1; o < 64; o <<= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { wv[wave] = bv; wi[wave] = bi; }
    __syncthreads();
    double pbv = wv[0];
    int p = wi[0];
    for (int w = 1; w < nwave; ++w)
      if (wv[w] > pbv || (wv[w] == pbv && wi[w] < p)) { pbv = wv[w]; p = wi[w]; }
    // the pivot row goes to urow, the row it replaces to xrow
    if (i == p) {
#pragma unroll
      for (int c = 0; c < kLuNb; ++c) urow[c] = v[c];
    }
    if (i == j && p != j) {
#pragma unroll
      for (int c = 0; c < kLuNb; ++c) xrow[c] = v[c];
    }
    if (t == 0) piv[jj] = p;
    __syncthreads();
    if (p != j && (i == j || i == p)) {
      const double* src = (i == j) ? urow : xrow;
#pragma unroll
      for (int c = 0; c < kLuNb; ++c) v[c] = src[c];
    }
    const double d = urow[0];
    if (d == 0.0) {  // exactly singular column (as DGETF2: record it, no elimination)
      if (t == 0 && info) atomicCAS(info + b, 0, (int)(j + 1));
    } else if (mine && i > j) {
      const double l = v[0] / d;
      v[0] = l;
#pragma unroll
      for (int c = 1; c < kLuNb; ++c) v[c] = fma(-l, urow[c], v[c]);
    }
    if (mine) A[i + j * nb] = v[0];
#pragma unroll
    for (int c = 0; c + 1 < kLuNb; ++c) v[c] = v[c + 1];
    v[kLuNb - 1] = 0.0;
  }
  __syncthreads();
  // the panel's row interchanges: on the panel column stored at step cc only
  // the later interchanges act; on the columns outside the panel all of them
  for (int64_t c = t; c < nb; c += blockDim.x) {
    const int first = (c >= k0 && c < k0 + jb) ? (int)(c - k0) + 1 : 0;
    for (int jj = first; jj < jb; ++jj) {
      const int64_t j = k0 + jj, p = piv[jj];
      if (p != j) {
        const double a = A[j + c * nb];
        A[j + c * nb] = A[p + c * nb];
        A[p + c * nb] = a;
      }
    }
  }
  __syncthreads();
  // inverse of the unit-lower diagonal block L11 (rows k0 .. k0+jb-1)
  for (int e = t; e < kLuNb * kLuNb; e += blockDim.x) {
    const int r = e % kLuNb, c = e / kLuNb;
    L[r][c] = (r < jb && c < r) ? A[(k0 + r) + (k0 + c) * nb] : 0.0;
  }
  __syncthreads();
  double* Li = Linv + (int64_t)b * kLuNb * kLuNb;
  if (t < kLuNb) {
    // column c of L11^{-1} by forward substitution in LDS; rows r < c are
    // zero and rows / columns >= jb give the identity (a full kLuNb block)
    const int c = t;
    for (int r = 0; r < kLuNb; ++r) Xi[r][c] = (r == c) ? 1.0 : 0.0;
    for (int r = c + 1; r < kLuNb; ++r) {
      double s = 0.0;
      for (int k = c; k < r; ++k) s = fma(L[r][k], Xi[k][c], s);
      Xi[r][c] = -s;
    }
    for (int r = 0; r < kLuNb; ++r) Li[r + (int64_t)c * kLuNb] = Xi[r][c];
  }
}

__global__ __launch_bounds__(kT) void copy_block_kernel(const double* S, int64_t lds, int64_t sS, double* D,
                                                        int64_t ldd, int64_t sD, int64_t m, int64_t n) {
  const double* Sb = S + blockIdx.y * sS;
  double* Db = D + blockIdx.y * sD;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < m * n; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m, j = e / m;
    Db[i + j * ldd] = Sb[i + j * lds];
  }
}

// ---------------------------------------------------------------------------
// dlacn2 on the LU of matrix b (order n = mats[idx[b]].n, leading dimension nb).
// ---------------------------------------------------------------------------
__device__ inline void solve_unit_lower(const double* W, int64_t ld, int n, double* x, int trans) {
  // trans = 0: L x = b (forward).  trans = 1: L^T x = b (backward).
  for (int s = 0; s < n; ++s) {
    const int j = trans ? n - 1 - s : s;
    __syncthreads();
    const double xj = x[j];
    if (!trans) {
      for (int r = j + 1 + threadIdx.x; r < n; r += kT) x[r] -= W[r + (int64_t)j * ld] * xj;
    } else {
      for (int r = threadIdx.x; r < j; r += kT) x[r] -= W[j + (int64_t)r * ld] * xj;
    }
  }
  __syncthreads();
}

__device__ inline void solve_upper(const double* W, int64_t ld, int n, double* x, int trans) {
  // trans = 0: U x = b (backward).  trans = 1: U^T x = b (forward).
  for (int s = 0; s < n; ++s) {
    const int j = trans ? s : n - 1 - s;
    __syncthreads();
    if (threadIdx.x == 0) x[j] /= W[j + (int64_t)j * ld];
    __syncthreads();
    const double xj = x[j];
    if (!trans) {
      for (int r = threadIdx.x; r < j; r += kT) x[r] -= W[r + (int64_t)j * ld] * xj;
    } else {
      for (int r = j + 1 + threadIdx.x; r < n; r += kT) x[r] -= W[j + (int64_t)r * ld] * xj;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kT) void condest_kernel(const BatchMat* mats, const int* idx, const double* W,
                                                     int64_t nb, const int* info, double* beta) {
  __shared__ double x[kMaxBatchN];
  __shared__ double sgn[kMaxBatchN];
  __shared__ double sv[kT];
  __shared__ int si[kT];
  __shared__ double shd[4];
  __shared__ int flag;
  const int b = blockIdx.x;
  const int g = idx[b];
  const int n = (int)mats[g].n;
  const int t = threadIdx.x;
  const double* Wb = W + (int64_t)b * nb * nb;
  if (info && info[b] != 0) {  // exactly singular: no positive lower bound
    if (t == 0) beta[g] = 0.0;
    return;
  }
  auto solve = [&](int trans) {
    if (!trans) {
      solve_unit_lower(Wb, nb, n, x, 0);
      solve_upper(Wb, nb, n, x, 0);
    } else {
      solve_upper(Wb, nb, n, x, 1);
      solve_unit_lower(Wb, nb, n, x, 1);
    }
  };
  auto asum = [&]() {
    double s = 0.0;
    for (int r = t; r < n; r += kT) s += fabs(x[r]);
    return block_sum256(s, shd);
  };
  auto iamax = [&]() {
    double bv = -1.0;
    int bi = n;
    for (int r = t; r < n; r += kT) {
      const double v = fabs(x[r]);
      if (v > bv) { bv = v; bi = r; }
    }
    sv[t] = bv;
    si[t] = bi;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
      if (t < s && (sv[t + s] > sv[t] || (sv[t + s] == sv[t] && si[t + s] < si[t]))) { sv[t] = sv[t + s]; si[t] = si[t + s]; }
      __syncthreads();
    }
    const int r = si[0];
    __syncthreads();
    return r;
  };
  for (int r = t; r < n; r += kT) x[r] = 1.0 / n;
  solve(0);
  double est;
  if (n == 1) {
    est = fabs(x[0]);
  } else {
    est = asum();
    for (int r = t; r < n; r += kT) {
      sgn[r] = x[r] >= 0 ? 1.0 : -1.0;
      x[r] = sgn[r];
    }
    solve(1);
    int j = iamax();
    int iter = 2;
    while (true) {
      for (int r = t; r < n; r += kT) x[r] = (r == j) ? 1.0 : 0.0;
      solve(0);
      const double estold = est;
      est = asum();
      if (t == 0) flag = 1;
      __syncthreads();
      for (int r = t; r < n; r += kT)
        if ((x[r] >= 0 ? 1.0 : -1.0) != sgn[r]) flag = 0;
      __syncthreads();
      const bool same = flag != 0;
      __syncthreads();
      if (same || est <= estold) break;
      for (int r = t; r < n; r += kT) {
        sgn[r] = x[r] >= 0 ? 1.0 : -1.0;
        x[r] = sgn[r];
      }
      solve(1);
      const int jlast = j;
      j = iamax();
      if (x[jlast] != fabs(x[j]) && iter < 5) {
        ++iter;
        continue;
      }
      break;
    }
    for (int r = t; r < n; r += kT) x[r] = ((r % 2 == 0) ? 1.0 : -1.0) * (1.0 + (double)r / (double)(n - 1));
    solve(0);
    const double temp = 2.0 * asum() / (3.0 * n);
    if (temp > est) est = temp;
  }
  if (t == 0) beta[g] = (est > 0.0 && isfinite(est)) ? 1.0 / (est * sqrt((double)n)) : 0.0;
}

// ---------------------------------------------------------------------------
// Zolotarev iterations (FP32, bucket layout: matrix b at X + b * nb * nb).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kT) void pad_scale_kernel(const BatchMat* mats, const int* idx, const double* alpha,
                                                       const double* beta, T* X, int64_t nb) {
  const int b = blockIdx.y;
  const int g = idx[b];
  const BatchMat m = mats[g];
  T* Xb = X + (int64_t)b * nb * nb;
  const double sc = 1.0 / (alpha[g] * beta[g]);
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < nb * nb; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % nb, j = e / nb;
    T v;
    if (i < m.n && j < m.n) v = (T)(sc * (double)m.A[i + j * m.lda]);
    else v = (i == j) ? (T)1 : (T)0;
    Xb[e] = v;
  }
}

// S_k = [X_{map[k]}; sdiag_k I]  (2nb x nb)
template <typename T>
__global__ __launch_bounds__(kT) void build_stack_items_kernel(const T* X, int64_t nb, const int* map,
                                                               const double* sdiag, T* S) {
  const int k = blockIdx.y;
  const T* Xb = X + (int64_t)map[k] * nb * nb;
  T* Sk = S + (int64_t)k * 2 * nb * nb;
  const T sd = (T)sdiag[k];
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < 2 * nb * nb; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % (2 * nb), j = e / (2 * nb);
    Sk[e] = (i < nb) ? Xb[i + j * nb] : ((i - nb == j) ? sd : (T)0);
  }
}

// Wq_k = lower(W_{map[k]}) + shift_k I, strict upper triangle zero
__global__ __launch_bounds__(kT) void shift_copy_items_kernel(const float* W, int64_t nb, const int* map,
                                                              const double* shift, float* Wq) {
  const int k = blockIdx.y;
  const float* Wb = W + (int64_t)map[k] * nb * nb;
  float* D = Wq + (int64_t)k * nb * nb;
  const double c = shift[k];
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < nb * nb; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % nb, j = e / nb;
    float v = 0.0f;
    if (i > j) v = Wb[e];
    else if (i == j) v = (float)((double)Wb[e] + c);
    D[e] = v;
  }
}

// X_b <- (X_b + sum_{k < zcnt[b]} Z[zoff[b] + k]) * ifone[b]; part[b][chunk] = sum (X+ - X)^2
__global__ __launch_bounds__(kT) void formx2_items_kernel(float* X, const float* Z, int64_t nn, const int* zoff,
                                                          const int* zcnt, const double* ifone, double* part) {
  __shared__ double sh[4];
  const int b = blockIdx.y;
  const int cnt = zcnt[b];
  double acc = 0.0;
  if (cnt > 0) {
    float* Xb = X + (int64_t)b * nn;
    const float* Zb = Z + (int64_t)zoff[b] * nn;
    const double f = ifone[b];
    for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < nn; e += (int64_t)gridDim.x * kT) {
      const float x = Xb[e];
      double s = (double)x;
      for (int k = 0; k < cnt; ++k) s += (double)Zb[(int64_t)k * nn + e];
      const float xp = (float)(s * f);
      const double d = (double)xp - (double)x;
      acc += d * d;
      Xb[e] = xp;
    }
  }
  acc = block_sum256(acc, sh);
  if (threadIdx.x == 0) part[(int64_t)b * gridDim.x + blockIdx.x] = acc;
}

// FP64 step of slot c (matrix loc[c] of the bucket): X32_loc <- (X64_c + sum Z64[zoff[c] + k]) * ifone[c],
// rounded to FP32; part[c][chunk] = sum (X+ - X)^2 in FP64
__global__ __launch_bounds__(kT) void formx2_f64_kernel(const double* X64, const double* Z, float* X32, int64_t nn,
                                                        const int* loc, const int* zoff, const int* zcnt,
                                                        const double* ifone, double* part) {
  __shared__ double sh[4];
  const int c = blockIdx.y;
  const int cnt = zcnt[c];
  const double* Xc = X64 + (int64_t)c * nn;
  const double* Zc = Z + (int64_t)zoff[c] * nn;
  float* Xo = X32 + (int64_t)loc[c] * nn;
  const double f = ifone[c];
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < nn; e += (int64_t)gridDim.x * kT) {
    const double x = Xc[e];
    double s = x;
    for (int k = 0; k < cnt; ++k) s += Zc[(int64_t)k * nn + e];
    const double xp = s * f;
    const double d = xp - x;
    acc += d * d;
    Xo[e] = (float)xp;
  }
  acc = block_sum256(acc, sh);
  if (threadIdx.x == 0) part[(int64_t)c * gridDim.x + blockIdx.x] = acc;
}

// out[gid[b]] = sum of the chunk partials of matrix b (matrices with zcnt = 0 untouched)
__global__ __launch_bounds__(kT) void sum_parts_kernel(const double* part, int nchunk, int count, const int* zcnt,
                                                       const int* gid, double* out) {
  const int b = blockIdx.x * kT + threadIdx.x;
  if (b >= count || zcnt[b] == 0) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)b * nchunk + c];
  out[gid[b]] = s;
}

__global__ __launch_bounds__(kT) void f2d_kernel(const float* X, double* Y, int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) Y[e] = (double)X[e];
}

// out[g] = ||G_b - I||_F^2 over the leading n x n block (lower triangle of G valid)
__global__ __launch_bounds__(kT) void gram_err_kernel(const double* G, int64_t nb, const BatchMat* mats,
                                                      const int* idx, double* out) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  const int g = idx[b];
  const double* Gb = G + (int64_t)b * nb * nb;
  const int64_t n = mats[g].n;
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n * n; e += kT) {
    const int64_t r = e % n, c = e / n;
    if (r < c) continue;
    const double v = Gb[r + c * nb] - (r == c ? 1.0 : 0.0);
    s += (r == c) ? v * v : 2.0 * v * v;
  }
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) out[g] = s;
}

__global__ __launch_bounds__(kT) void fill_int_kernel(int* p, int count, int value) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < count) p[i] = value;
}

__global__ __launch_bounds__(kT) void unpad_kernel(const float* X, int64_t nb, const BatchMat* mats, const int* idx) {
  const int b = blockIdx.y;
  const BatchMat m = mats[idx[b]];
  const float* Xb = X + (int64_t)b * nb * nb;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < m.n * m.n; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m.n, j = e / m.n;
    m.U[i + j * m.ldu] = Xb[i + j * nb];
  }
}

}  // namespace

int kb_fill_int(int* p, int count, int value, hipStream_t s) {
  if (count <= 0) return kOk;
  hipLaunchKernelGGL(fill_int_kernel, dim3((count + kT - 1) / kT), dim3(kT), 0, s, p, count, value);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_alpha(const BatchMat* mats, int count, double* alpha, hipStream_t s) {
  hipLaunchKernelGGL(alpha_kernel, dim3(count), dim3(kT), 0, s, mats, alpha);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_lu_pad(const BatchMat* mats, const int* idx, const double* alpha, double* W, int64_t nb, int count,
              hipStream_t s) {
  hipLaunchKernelGGL(lu_pad_kernel, dim3(gfor(nb * nb, 64), count), dim3(kT), 0, s, mats, idx, alpha, W, nb);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_lu_panel(double* W, int64_t nb, int64_t k0, double* Linv, int* info, int count, hipStream_t s) {
  if (nb - k0 > kLuRows || nb - k0 <= 0) {
    set_error("batched LU panel: size out of range");
    return kErrArg;
  }
  const unsigned threads = (unsigned)((nb - k0 + 63) / 64 * 64);
  hipLaunchKernelGGL(lu_panel_wg_kernel, dim3(count), dim3(threads), 0, s, W, nb, k0, Linv, info);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_copy_block(const double* S, int64_t lds, int64_t sS, double* D, int64_t ldd, int64_t sD, int64_t m, int64_t n,
                  int count, hipStream_t s) {
  if (m <= 0 || n <= 0) return kOk;
  hipLaunchKernelGGL(copy_block_kernel, dim3(gfor(m * n, 64), count), dim3(kT), 0, s, S, lds, sS, D, ldd, sD, m, n);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_condest(const BatchMat* mats, const int* idx, const double* W, int64_t nb, const int* info, double* beta,
               int count, hipStream_t s) {
  hipLaunchKernelGGL(condest_kernel, dim3(count), dim3(kT), 0, s, mats, idx, W, nb, info, beta);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int kb_pad_scale(const BatchMat* mats, const int* idx, const double* alpha, const double* beta, T* X, int64_t nb,
                 int count, hipStream_t s) {
  hipLaunchKernelGGL((pad_scale_kernel<T>), dim3(gfor(nb * nb, 64), count), dim3(kT), 0, s, mats, idx, alpha, beta, X,
                     nb);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}
template int kb_pad_scale<float>(const BatchMat*, const int*, const double*, const double*, float*, int64_t, int,
                                 hipStream_t);
template int kb_pad_scale<double>(const BatchMat*, const int*, const double*, const double*, double*, int64_t, int,
                                  hipStream_t);

template <typename T>
int kb_build_stack(const T* X, int64_t nb, const int* map, const double* sdiag, T* S, int items, hipStream_t s) {
  hipLaunchKernelGGL((build_stack_items_kernel<T>), dim3(gfor(2 * nb * nb, 64), items), dim3(kT), 0, s, X, nb, map,
                     sdiag, S);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}
template int kb_build_stack<float>(const float*, int64_t, const int*, const double*, float*, int, hipStream_t);
template int kb_build_stack<double>(const double*, int64_t, const int*, const double*, double*, int, hipStream_t);

int kb_formx2_f64(const double* X64, const double* Z, float* X32, int64_t nb, const int* loc, const int* zoff,
                  const int* zcnt, const double* ifone, double* part, int nchunk, const int* gid, double* err2,
                  int count, hipStream_t s) {
  hipLaunchKernelGGL(formx2_f64_kernel, dim3(nchunk, count), dim3(kT), 0, s, X64, Z, X32, nb * nb, loc, zoff, zcnt,
                     ifone, part);
  hipLaunchKernelGGL(sum_parts_kernel, dim3((count + kT - 1) / kT), dim3(kT), 0, s, part, nchunk, count, zcnt, gid,
                     err2);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_shift_copy(const float* W, int64_t nb, const int* map, const double* shift, float* Wq, int items,
                  hipStream_t s) {
  hipLaunchKernelGGL(shift_copy_items_kernel, dim3(gfor(nb * nb, 64), items), dim3(kT), 0, s, W, nb, map, shift, Wq);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_formx2(float* X, const float* Z, int64_t nb, const int* zoff, const int* zcnt, const double* ifone,
              double* part, int nchunk, const int* gid, double* err2, int count, hipStream_t s) {
  hipLaunchKernelGGL(formx2_items_kernel, dim3(nchunk, count), dim3(kT), 0, s, X, Z, nb * nb, zoff, zcnt, ifone, part);
  hipLaunchKernelGGL(sum_parts_kernel, dim3((count + kT - 1) / kT), dim3(kT), 0, s, part, nchunk, count, zcnt, gid,
                     err2);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_f2d(const float* X, double* Y, int64_t total, hipStream_t s) {
  hipLaunchKernelGGL(f2d_kernel, dim3(gfor(total, 8192)), dim3(kT), 0, s, X, Y, total);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_gram_err(const double* G, int64_t nb, const BatchMat* mats, const int* idx, double* out, int count,
                hipStream_t s) {
  hipLaunchKernelGGL(gram_err_kernel, dim3(count), dim3(kT), 0, s, G, nb, mats, idx, out);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kb_unpad(const float* X, int64_t nb, const BatchMat* mats, const int* idx, int count, hipStream_t s) {
  hipLaunchKernelGGL(unpad_kernel, dim3(gfor(nb * nb, 64), count), dim3(kT), 0, s, X, nb, mats, idx);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

}  // namespace zolo
