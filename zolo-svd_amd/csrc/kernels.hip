// Memory-bound and small-tile kernels of the polar-decomposition path:
// fused FormX2 update, Frobenius / 1-norm reductions, stacked-matrix build,
// shifted Gram copies, the LDS-resident Cholesky+inverse leaf, symmetrisation
// and precision conversion.  All are batched over the Zolotarev shifts
// through blockIdx.y / blockIdx.z where a shift index applies.
#include "kernels.hpp"

namespace zolo {
namespace {

constexpr int kT = 256;

__device__ inline double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block-wide sum of a double; result valid in thread 0.
__device__ inline double block_sum(double v, double* sh) {
  v = warp_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// X+ = (X + sum_j Z_j) / fone ; partial[blk] = sum (X+ - X)^2     (FormX2)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kT) void formx2_kernel(T* X, int64_t ldx, const T* Z, int64_t ldz, int64_t strideZ,
                                                   int nz, int64_t n, double inv_fone, double* partial) {
  __shared__ double sh[8];
  const int64_t total = n * n;
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % n, j = e / n;
    const T x = X[i + j * ldx];
    double s = (double)x;
    for (int q = 0; q < nz; ++q) s += (double)Z[q * strideZ + i + j * ldz];
    const T xp = (T)(s * inv_fone);
    const double d = (double)xp - (double)x;
    acc += d * d;
    X[i + j * ldx] = xp;
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// Sum of squares of a matrix (partials per block); mode 1 = sum of squares of (A - B).
template <typename T>
__global__ __launch_bounds__(kT) void sumsq_kernel(const T* A, int64_t lda, int64_t m, int64_t n, double* partial) {
  __shared__ double sh[8];
  const int64_t total = m * n;
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m, j = e / m;
    const double v = (double)A[i + j * lda];
    acc += v * v;
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// Column abs sums -> colsum[j]; one block per column.
template <typename T>
__global__ __launch_bounds__(kT) void colabs_kernel(const T* A, int64_t lda, int64_t m, double* colsum) {
  __shared__ double sh[8];
  const int64_t j = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += kT) acc += fabs((double)A[i + j * lda]);
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) colsum[j] = r;
}

// Final deterministic reduction: out[0] = sum or max of v[0..n).
__global__ __launch_bounds__(kT) void reduce_kernel(const double* v, int64_t n, int is_max, double* out) {
  __shared__ double sh[kT];
  double acc = is_max ? 0.0 : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kT) acc = is_max ? fmax(acc, v[i]) : acc + v[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = is_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + s]) : sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// B = alpha * A (m x n), with optional precision conversion.
template <typename TA, typename TB>
__global__ __launch_bounds__(kT) void scale_copy_kernel(const TA* A, int64_t lda, TB* B, int64_t ldb, int64_t m, int64_t n,
                                                        double alpha) {
  const int64_t total = m * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m, j = e / m;
    B[i + j * ldb] = (TB)(alpha * (double)A[i + j * lda]);
  }
}

// C = a*A + b*B (m x n).
template <typename T>
__global__ __launch_bounds__(kT) void axpby_kernel(double a, const T* A, int64_t lda, double b, const T* B, int64_t ldb,
                                                   T* C, int64_t ldc, int64_t m, int64_t n) {
  const int64_t total = m * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m, j = e / m;
    C[i + j * ldc] = (T)(a * (double)A[i + j * lda] + b * (double)B[i + j * ldb]);
  }
}

// Stacked matrix S_q = [X; s_q I] (2n x n, ld = lds), batched over q = blockIdx.y.
template <typename T>
__global__ __launch_bounds__(kT) void build_stack_kernel(const T* X, int64_t ldx, T* S, int64_t lds, int64_t strideS,
                                                         int64_t n, const double* sdiag, int upper) {
  const int q = blockIdx.y;
  T* Sq = S + q * strideS;
  const double s = sdiag[q];
  const int64_t total = 2 * n * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % (2 * n), j = e / (2 * n);
    T v;
    if (i < n) v = (upper && i > j) ? (T)0 : X[i + j * ldx];
    else v = (i - n == j) ? (T)s : (T)0;
    Sq[i + j * lds] = v;
  }
}

// W_q = lower(W) + c_q I, batched over q; the strict upper triangle is zeroed.
template <typename T>
__global__ __launch_bounds__(kT) void shift_copy_kernel(const T* W, int64_t ldw, T* Wq, int64_t ldq, int64_t strideQ,
                                                        int64_t n, const double* shift, double scale, double add) {
  const int q = blockIdx.y;
  T* D = Wq + q * strideQ;
  const double c = shift ? shift[q] : 0.0;
  const int64_t total = n * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % n, j = e / n;
    T v = (T)0;
    if (i >= j) {
      v = (T)(scale * (double)W[i + j * ldw]);
      if (i == j) v = (T)((double)v + c + add);
    }
    D[i + j * ldq] = v;
  }
}

// A = value (m x n, batched over blockIdx.y).
template <typename T>
__global__ __launch_bounds__(kT) void fill_kernel(T* A, int64_t lda, int64_t strideA, int64_t m, int64_t n, T value) {
  T* D = A + blockIdx.y * strideA;
  const int64_t total = m * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT)
    D[(e % m) + (e / m) * lda] = value;
}

// Zero the strict upper triangle (batched).
template <typename T>
__global__ __launch_bounds__(kT) void zero_upper_kernel(T* A, int64_t lda, int64_t strideA, int64_t n) {
  T* D = A + blockIdx.y * strideA;
  const int64_t total = n * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % n, j = e / n;
    if (i < j) D[i + j * lda] = (T)0;
  }
}

// A <- (A + A^T)/2 in place, both triangles.  One workgroup per pair of
// 32 x 32 tiles (bi, bj), bi >= bj (the others exit): both tiles are read
// along columns into LDS, and element (i, j) and (j, i) receive the same
// value 0.5 (a_ij + a_ji) (the sum commutes exactly).
constexpr int kSymT = 32;
template <typename T>
__global__ __launch_bounds__(256) void symmetrize_kernel(T* A, int64_t lda, int64_t n) {
  const int64_t bi = blockIdx.x, bj = blockIdx.y;
  if (bi < bj) return;
  __shared__ double t1[kSymT][kSymT + 1], t2[kSymT][kSymT + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = bi * kSymT, c0 = bj * kSymT;
  // t1[x][k] = A(r0 + x, c0 + k), t2[x][k] = A(c0 + x, r0 + k)
  for (int k = ty; k < kSymT; k += 8) {
    t1[tx][k] = (r0 + tx < n && c0 + k < n) ? (double)A[(r0 + tx) + (c0 + k) * lda] : 0.0;
    t2[tx][k] = (c0 + tx < n && r0 + k < n) ? (double)A[(c0 + tx) + (r0 + k) * lda] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < kSymT; k += 8) {
    const int64_t r = r0 + tx, c = c0 + k;  // A(r, c) = (A(r, c) + A(c, r)) / 2
    if (r < n && c < n) A[r + c * lda] = (T)(0.5 * (t1[tx][k] + t2[k][tx]));
  }
  if (bi == bj) return;  // the diagonal tile is complete
  for (int k = ty; k < kSymT; k += 8) {
    const int64_t r = c0 + tx, c = r0 + k;  // A(r, c) = (A(c, r) + A(r, c)) / 2
    if (r < n && c < n) A[r + c * lda] = (T)(0.5 * (t1[k][tx] + t2[tx][k]));
  }
}

// A += diag(d) ; d scalar
template <typename T>
__global__ __launch_bounds__(kT) void add_diag_kernel(T* A, int64_t lda, int64_t n, double d) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    A[i + i * lda] = (T)((double)A[i + i * lda] + d);
}

// A(i, j) *= f for i <= j (the upper triangle; the rest is left alone)
__global__ __launch_bounds__(kT) void scale_upper_kernel(double* A, int64_t lda, int64_t n, double f) {
  const int64_t total = n * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % n, j = e / n;
    if (i <= j) A[i + j * lda] *= f;
  }
}

// A(ri[e], ci[e]) = v[e]; positions distinct (no two threads store to one)
__global__ __launch_bounds__(kT) void scatter_kernel(double* A, int64_t lda, const int64_t* ri, const int64_t* ci,
                                                     const double* v, int64_t z) {
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < z; e += (int64_t)gridDim.x * kT)
    A[ri[e] + ci[e] * lda] = v[e];
}

// Symmetric MatrixMarket assembly: A <- A + A^T off the diagonal and
// A(i, i) <- (A(i, i) + A(i, i))/2 - sigma on it.  One 32 x 32 tile pair per
// workgroup (tile row bi >= tile column bj; the others exit): both tiles are
// read and written along columns through LDS.
constexpr int kSymTile = 32;
__global__ __launch_bounds__(256) void mm_symmetrize_kernel(double* A, int64_t lda, int64_t n, double sigma) {
  const int64_t bi = blockIdx.x, bj = blockIdx.y;
  if (bi < bj) return;
  __shared__ double t1[kSymTile][kSymTile + 1], t2[kSymTile][kSymTile + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = bi * kSymTile, c0 = bj * kSymTile;
  // t1[x][k] = A(r0 + x, c0 + k), t2[x][k] = A(c0 + x, r0 + k)
  for (int k = ty; k < kSymTile; k += 8) {
    t1[tx][k] = (r0 + tx < n && c0 + k < n) ? A[(r0 + tx) + (c0 + k) * lda] : 0.0;
    t2[tx][k] = (c0 + tx < n && r0 + k < n) ? A[(c0 + tx) + (r0 + k) * lda] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < kSymTile; k += 8) {
    const int64_t r = r0 + tx, c = c0 + k;
    if (r < n && c < n) {  // A(r, c) + A(c, r)
      const double v = (r == c) ? (t1[tx][k] + t1[tx][k]) / 2.0 - sigma : t1[tx][k] + t2[k][tx];
      A[r + c * lda] = v;
    }
  }
  if (bi == bj) return;  // the diagonal tile is complete
  for (int k = ty; k < kSymTile; k += 8) {
    const int64_t r = c0 + tx, c = r0 + k;
    if (r < n && c < n) A[r + c * lda] = t1[k][tx] + t2[tx][k];  // A(c, r) + A(r, c)
  }
}

// ---------------------------------------------------------------------------
// Cholesky + triangular inverse of one nb x nb diagonal block in LDS.
// In: lower triangle of A (n x n block at A, ld lda).  Out: L in A (strict upper
// zeroed), L^{-1} in Li (strict upper zeroed).  Batched over blockIdx.x.
// info[q] receives (global column + 1) of the first non-positive pivot.
// ---------------------------------------------------------------------------
// TPB threads per block, G = TPB / 64 column groups.  Four waves: a single
// wave needs no s_barrier (its LDS accesses complete in order) but then runs
// every row's update serially, 224 instead of 95 us per 64 x 64 block.
constexpr int kLeafTPB = 256;
template <typename T, int NB, int TPB>
__global__ __launch_bounds__(TPB) void chol_leaf_kernel(T* A, int64_t lda, int64_t strideA, T* Li, int64_t ldl,
                                                        int64_t strideL, int nb, int64_t col0, int* info) {
  __shared__ T a[NB][NB + 1];
  __shared__ T li[NB][NB + 1];
  __shared__ int bad;
  constexpr int G = TPB / 64;  // column groups
  static_assert(NB == 64 && TPB % 64 == 0 && (G & (G - 1)) == 0, "one row per lane, G column groups");
  const int q = blockIdx.x;
  T* Aq = A + q * strideA;
  T* Lq = Li + q * strideL;
  const int t = threadIdx.x;
  // thread (r, g): row r of the block, columns k = g (mod G) of it
  const int r = t & 63, g = t >> 6;
  if (t == 0) bad = 0;
  for (int e = t; e < nb * nb; e += TPB) {
    const int i = e % nb, j = e / nb;
    a[i][j] = (i >= j) ? Aq[i + j * lda] : (T)0;
    li[i][j] = (T)0;
  }
  __syncthreads();
  // right-looking Cholesky: per column the pivot, the scaled column, the
  // rank-1 update of the lower trailing block (row r, columns k = g mod G)
  for (int j = 0; j < nb; ++j) {
    if (t == 0) {
      const T d = a[j][j];
      if (!(d > (T)0)) {
        if (!bad) bad = j + 1;
        a[j][j] = (T)1;  // keep going so every wave reaches the end
      } else {
        a[j][j] = sqrt(d);
      }
    }
    __syncthreads();
    if (g == 0 && r > j && r < nb) a[r][j] /= a[j][j];
    __syncthreads();
    if (r > j && r < nb) {
      const T arj = a[r][j];
      for (int k = j + 1 + ((g - j - 1) & (G - 1)); k <= r; k += G) a[r][k] -= arj * a[k][j];
    }
    __syncthreads();
  }
  // Inverse by forward substitution L x = e_c for all columns at once (axpy
  // form): at step k, row k of the inverse is divided by l_kk (it is then
  // final) and updates the rows below; each element takes the same terms in
  // the same order as a per-column substitution.
  if (t < nb) li[t][t] = (T)1;
  __syncthreads();
  for (int k = 0; k < nb; ++k) {
    if (t <= k) li[k][t] /= a[k][k];
    __syncthreads();
    if (r > k && r < nb) {
      const T ark = a[r][k];
      for (int c = g; c <= k; c += G) li[r][c] -= ark * li[k][c];
    }
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += TPB) {
    const int i = e % nb, j = e / nb;
    Aq[i + j * lda] = a[i][j];
    Lq[i + j * ldl] = li[i][j];
  }
  if (t == 0 && bad && info) atomicCAS(&info[q], 0, (int)(col0 + bad));
}

// Uniform [0, 1) entries from a 64-bit hash of (seed, column-major index):
// the SplitMix64 finaliser (S. Vigna), top 53 bits as the mantissa.  The
// random start block of the subspace iteration (pqdwhsubit.f90).
__global__ __launch_bounds__(kT) void fill_uniform_kernel(double* X, int64_t ld, int64_t m, int64_t n,
                                                          unsigned long long seed) {
  const int64_t total = m * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % m, j = e / m;
    unsigned long long z = seed + (unsigned long long)(e + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    X[i + j * ld] = (double)(z >> 11) * 0x1.0p-53;
  }
}

inline unsigned grid_for(int64_t total) {
  int64_t g = (total + kT - 1) / kT;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename T>
int k_formx2(T* X, int64_t ldx, const T* Z, int64_t ldz, int64_t strideZ, int nz, int64_t n, double fone,
             double* partial, int npartial, double* out, hipStream_t s) {
  hipLaunchKernelGGL((formx2_kernel<T>), dim3(npartial), dim3(kT), 0, s, X, ldx, Z, ldz, strideZ, nz, n, 1.0 / fone,
                     partial);
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kT), 0, s, partial, (int64_t)npartial, 0, out);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_sumsq(const T* A, int64_t lda, int64_t m, int64_t n, double* partial, int npartial, double* out, hipStream_t s) {
  hipLaunchKernelGGL((sumsq_kernel<T>), dim3(npartial), dim3(kT), 0, s, A, lda, m, n, partial);
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kT), 0, s, partial, (int64_t)npartial, 0, out);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_norm1(const T* A, int64_t lda, int64_t m, int64_t n, double* colsum, double* out, hipStream_t s) {
  hipLaunchKernelGGL((colabs_kernel<T>), dim3((unsigned)n), dim3(kT), 0, s, A, lda, m, colsum);
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kT), 0, s, colsum, n, 1, out);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename TA, typename TB>
int k_scale_copy(const TA* A, int64_t lda, TB* B, int64_t ldb, int64_t m, int64_t n, double alpha, hipStream_t s) {
  hipLaunchKernelGGL((scale_copy_kernel<TA, TB>), dim3(grid_for(m * n)), dim3(kT), 0, s, A, lda, B, ldb, m, n, alpha);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_axpby(double a, const T* A, int64_t lda, double b, const T* B, int64_t ldb, T* C, int64_t ldc, int64_t m,
            int64_t n, hipStream_t s) {
  hipLaunchKernelGGL((axpby_kernel<T>), dim3(grid_for(m * n)), dim3(kT), 0, s, a, A, lda, b, B, ldb, C, ldc, m, n);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_build_stack(const T* X, int64_t ldx, T* S, int64_t lds, int64_t strideS, int64_t n, const double* sdiag,
                  int batch, hipStream_t s, bool upper) {
  hipLaunchKernelGGL((build_stack_kernel<T>), dim3(grid_for(2 * n * n), batch), dim3(kT), 0, s, X, ldx, S, lds, strideS,
                     n, sdiag, upper ? 1 : 0);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_shift_copy(const T* W, int64_t ldw, T* Wq, int64_t ldq, int64_t strideQ, int64_t n, const double* shift,
                 double scale, double add, int batch, hipStream_t s) {
  hipLaunchKernelGGL((shift_copy_kernel<T>), dim3(grid_for(n * n), batch), dim3(kT), 0, s, W, ldw, Wq, ldq, strideQ, n,
                     shift, scale, add);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_fill(T* A, int64_t lda, int64_t strideA, int64_t m, int64_t n, double value, int batch, hipStream_t s) {
  hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(m * n), batch), dim3(kT), 0, s, A, lda, strideA, m, n, (T)value);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_zero_upper(T* A, int64_t lda, int64_t strideA, int64_t n, int batch, hipStream_t s) {
  hipLaunchKernelGGL((zero_upper_kernel<T>), dim3(grid_for(n * n), batch), dim3(kT), 0, s, A, lda, strideA, n);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_symmetrize(T* A, int64_t lda, int64_t n, hipStream_t s) {
  if (n <= 0) return kOk;
  const int64_t nt = (n + kSymT - 1) / kSymT;
  if (nt > 65535) {
    set_error("symmetrize: order too large");
    return kErrArg;
  }
  hipLaunchKernelGGL((symmetrize_kernel<T>), dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, s, A, lda, n);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_add_diag(T* A, int64_t lda, int64_t n, double d, hipStream_t s) {
  hipLaunchKernelGGL((add_diag_kernel<T>), dim3(grid_for(n)), dim3(kT), 0, s, A, lda, n, d);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

template <typename T>
int k_chol_leaf(T* A, int64_t lda, int64_t strideA, T* Li, int64_t ldl, int64_t strideL, int nb, int64_t col0,
                int* info, int batch, hipStream_t s) {
  if (nb > kCholLeaf || nb <= 0) {
    set_error("chol leaf: block size out of range");
    return kErrArg;
  }
  hipLaunchKernelGGL((chol_leaf_kernel<T, kCholLeaf, kLeafTPB>), dim3(batch), dim3(kLeafTPB), 0, s, A, lda, strideA, Li,
                     ldl, strideL, nb, col0, info);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int k_scale_upper(double* A, int64_t lda, int64_t n, double f, hipStream_t s) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(scale_upper_kernel, dim3(grid_for(n * n)), dim3(kT), 0, s, A, lda, n, f);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int k_scatter(double* A, int64_t lda, const int64_t* ri, const int64_t* ci, const double* v, int64_t z,
              hipStream_t s) {
  if (z <= 0) return kOk;
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(z)), dim3(kT), 0, s, A, lda, ri, ci, v, z);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int k_mm_symmetrize(double* A, int64_t lda, int64_t n, double sigma, hipStream_t s) {
  if (n <= 0) return kOk;
  const int64_t nt = (n + kSymTile - 1) / kSymTile;
  if (nt > 65535) {
    set_error("symmetric assembly: order too large");
    return kErrArg;
  }
  hipLaunchKernelGGL(mm_symmetrize_kernel, dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, s, A, lda, n, sigma);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int k_fill_uniform(double* X, int64_t ld, int64_t m, int64_t n, unsigned long long seed, hipStream_t s) {
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(m * n)), dim3(kT), 0, s, X, ld, m, n, seed);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

#define ZOLO_INST(T)                                                                                                    \
  template int k_formx2<T>(T*, int64_t, const T*, int64_t, int64_t, int, int64_t, double, double*, int, double*,        \
                           hipStream_t);                                                                                \
  template int k_sumsq<T>(const T*, int64_t, int64_t, int64_t, double*, int, double*, hipStream_t);                      \
  template int k_norm1<T>(const T*, int64_t, int64_t, int64_t, double*, double*, hipStream_t);                           \
  template int k_axpby<T>(double, const T*, int64_t, double, const T*, int64_t, T*, int64_t, int64_t, int64_t,         \
                          hipStream_t);                                                                                 \
  template int k_build_stack<T>(const T*, int64_t, T*, int64_t, int64_t, int64_t, const double*, int, hipStream_t, bool);     \
  template int k_shift_copy<T>(const T*, int64_t, T*, int64_t, int64_t, int64_t, const double*, double, double, int,    \
                               hipStream_t);                                                                            \
  template int k_zero_upper<T>(T*, int64_t, int64_t, int64_t, int, hipStream_t);                                        \
  template int k_fill<T>(T*, int64_t, int64_t, int64_t, int64_t, double, int, hipStream_t);                             \
  template int k_symmetrize<T>(T*, int64_t, int64_t, hipStream_t);                                                      \
  template int k_add_diag<T>(T*, int64_t, int64_t, double, hipStream_t);                                                \
  template int k_chol_leaf<T>(T*, int64_t, int64_t, T*, int64_t, int64_t, int, int64_t, int*, int, hipStream_t);
ZOLO_INST(double)
ZOLO_INST(float)
template int k_scale_copy<double, double>(const double*, int64_t, double*, int64_t, int64_t, int64_t, double, hipStream_t);
template int k_scale_copy<double, float>(const double*, int64_t, float*, int64_t, int64_t, int64_t, double, hipStream_t);
template int k_scale_copy<float, double>(const float*, int64_t, double*, int64_t, int64_t, int64_t, double, hipStream_t);
template int k_scale_copy<float, float>(const float*, int64_t, float*, int64_t, int64_t, int64_t, double, hipStream_t);

}  // namespace zolo
