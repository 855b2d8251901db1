// FP64 polish of a polar factor computed with low-precision Cholesky steps.
//
// The low-precision iterations return U ~ polar(Xref) with an error of the
// size of their unit roundoff delta, where Xref is the last FP64 iterate (the
// output of the FP64 QR step; it has A's polar factor to FP64 accuracy and is
// well conditioned).  One pass computes, in FP64,
//     F = U^T U - I   and   C = U^T Xref,
// splits C = Hs + K (symmetric + skew), and solves in low precision the
// Lyapunov equation  S Hs + Hs S = 2K - (F Hs - Hs F)/2  for a skew S by
// conjugate gradients (the operator has eigenvalues h_i + h_j, so it is as
// well conditioned as Xref).  U <- U (I + S - F/2) then makes U^T U = I and
// U^T Xref symmetric to first order, so each pass squares the error.  A pass
// costs one FP64 SYRK-shaped GEMM (n^3) and one FP64 GEMM (2 n^3); the CG and
// the update GEMM run in the low precision.
#include <algorithm>
#include <cmath>

#include "driver.hpp"

namespace zolo {
namespace {

constexpr int kT = 256;
constexpr int kNp = 1024;

inline unsigned grid_for(int64_t total) {
  int64_t g = (total + kT - 1) / kT;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

// Elementwise functions of (X[i][j], X[j][i]) over 64 x 64 tiles: the tile of
// X^T is read in coalesced columns into LDS and read back transposed.
//   MODE 0: out1 = F = lower(X) mirrored minus I (X holds U^T U in its lower triangle)
//   MODE 1: out1 = Hs = (X + X^T)/2, out2 = R = X - X^T (= 2K)
template <int MODE>
__global__ __launch_bounds__(kT) void pair_tile_kernel(const double* X, int64_t n, float* out1, float* out2) {
  __shared__ double tt[64][65];
  const int64_t bi = (int64_t)blockIdx.x * 64, bj = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (MODE == 1 || bi < bj + 64) {  // MODE 0 reads the transpose only above the diagonal
    for (int b = ty; b < 64; b += kT / 64)  // tt[b][a] = X[bj + a][bi + b]
      tt[b][tx] = (bj + tx < n && bi + b < n) ? X[(bj + tx) + (bi + b) * n] : 0.0;
  }
  __syncthreads();
  const int64_t i = bi + tx;
  if (i >= n) return;
  for (int r = ty; r < 64; r += kT / 64) {
    const int64_t j = bj + r;
    if (j >= n) break;
    const int64_t e = i + j * n;
    if (MODE == 0) {
      double v = (i >= j) ? X[e] : tt[tx][r];
      if (i == j) v -= 1.0;
      out1[e] = (float)v;
    } else {
      const double a = X[e], b = tt[tx][r];
      out1[e] = (float)(0.5 * (a + b));
      out2[e] = (float)(a - b);
    }
  }
}

// 64 x 64 tiles of D = M - M^T (ACC = false) or D += coef (M - M^T) (ACC = true):
// the tile of M^T is read in coalesced columns into LDS and read back
// transposed (row pitch 65: no bank conflicts).
template <bool ACC>
__global__ __launch_bounds__(kT) void skew_tile_kernel(float* D, const float* M, int64_t n, float coef) {
  __shared__ float tt[64][65];
  const int64_t bi = (int64_t)blockIdx.x * 64, bj = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // tt[b][a] = M[bj + a][bi + b]
  for (int b = ty; b < 64; b += kT / 64)
    tt[b][tx] = (bj + tx < n && bi + b < n) ? M[(bj + tx) + (bi + b) * n] : 0.f;
  __syncthreads();
  if (bi + tx >= n) return;
  for (int r = ty; r < 64; r += kT / 64) {
    if (bj + r >= n) break;
    const int64_t e = (bi + tx) + (bj + r) * n;
    const float v = M[e] - tt[tx][r];
    if (ACC) D[e] += coef * v;
    else D[e] = v;
  }
}

// 64 x 64 tiles of D = M - M^T (the Lyapunov operator applied to the CG
// direction P, with M = P Hs) and, in the same pass, each tile's sum of
// P .* D in double: part[tile] (tiles numbered x-fastest; the sums run in a
// fixed order, so the result does not depend on scheduling).
__global__ __launch_bounds__(kT) void skew_dot_kernel(float* D, const float* M, const float* P, int64_t n,
                                                      double* part) {
  __shared__ float tt[64][65];
  __shared__ double sh[kT];
  const int64_t bi = (int64_t)blockIdx.x * 64, bj = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // tt[b][a] = M[bj + a][bi + b]
  for (int b = ty; b < 64; b += kT / 64)
    tt[b][tx] = (bj + tx < n && bi + b < n) ? M[(bj + tx) + (bi + b) * n] : 0.f;
  __syncthreads();
  double acc = 0.0;
  if (bi + tx < n) {
    for (int r = ty; r < 64; r += kT / 64) {
      if (bj + r >= n) break;
      const int64_t e = (bi + tx) + (bj + r) * n;
      const float v = M[e] - tt[tx][r];
      D[e] = v;
      acc += (double)P[e] * (double)v;
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kT / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// One CG step on the iterate and the residual, four floats per thread and
// iteration: S += a P, R -= a AP, and part[block] = the block's sum of R^2
// (double) of the updated residual.
__global__ __launch_bounds__(kT) void cg_update_kernel(float* __restrict__ S, float* __restrict__ R,
                                                       const float* __restrict__ P, const float* __restrict__ AP,
                                                       int64_t len, float a, double* part) {
  __shared__ double sh[kT];
  double acc = 0.0;
  const int64_t n4 = len / 4, stride = (int64_t)gridDim.x * kT;
  float4* S4 = reinterpret_cast<float4*>(S);
  float4* R4 = reinterpret_cast<float4*>(R);
  const float4* P4 = reinterpret_cast<const float4*>(P);
  const float4* A4 = reinterpret_cast<const float4*>(AP);
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < n4; e += stride) {
    const float4 p = P4[e], ap = A4[e];
    float4 sv = S4[e], r = R4[e];
    sv.x += a * p.x; sv.y += a * p.y; sv.z += a * p.z; sv.w += a * p.w;
    r.x -= a * ap.x; r.y -= a * ap.y; r.z -= a * ap.z; r.w -= a * ap.w;
    S4[e] = sv;
    R4[e] = r;
    acc += ((double)r.x * r.x + (double)r.y * r.y) + ((double)r.z * r.z + (double)r.w * r.w);
  }
  for (int64_t e = n4 * 4 + (int64_t)blockIdx.x * kT + threadIdx.x; e < len; e += stride) {
    S[e] += a * P[e];
    R[e] -= a * AP[e];
    acc += (double)R[e] * (double)R[e];
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kT / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// partial[blk] = sum a*b (double accumulation)
__global__ __launch_bounds__(kT) void dot_kernel(const float* a, const float* b, int64_t len, double* partial) {
  __shared__ double sh[kT];
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < len; e += (int64_t)gridDim.x * kT)
    s += (double)a[e] * (double)b[e];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = kT / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kT) void maxabs_kernel(const float* a, int64_t len, double* partial) {
  __shared__ double sh[kT];
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < len; e += (int64_t)gridDim.x * kT)
    s = fmax(s, fabs((double)a[e]));
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = kT / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kT) void finish_reduce_kernel(const double* p, int np, int is_max, double* out) {
  __shared__ double sh[kT];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += kT) s = is_max ? fmax(s, p[i]) : s + p[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = kT / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] = is_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + k]) : sh[threadIdx.x] + sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// p = r + b*p, four floats per thread and iteration
__global__ __launch_bounds__(kT) void xpby_f_kernel(float* __restrict__ p, const float* __restrict__ r, int64_t len,
                                                    float b) {
  const int64_t n4 = len / 4, stride = (int64_t)gridDim.x * kT;
  float4* P4 = reinterpret_cast<float4*>(p);
  const float4* R4 = reinterpret_cast<const float4*>(r);
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < n4; e += stride) {
    const float4 rv = R4[e];
    float4 pv = P4[e];
    pv.x = rv.x + b * pv.x; pv.y = rv.y + b * pv.y; pv.z = rv.z + b * pv.z; pv.w = rv.w + b * pv.w;
    P4[e] = pv;
  }
  for (int64_t e = n4 * 4 + (int64_t)blockIdx.x * kT + threadIdx.x; e < len; e += stride) p[e] = r[e] + b * p[e];
}
// E = S - F/2
__global__ __launch_bounds__(kT) void form_e_kernel(float* E, const float* S, const float* F, int64_t len) {
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < len; e += (int64_t)gridDim.x * kT) E[e] = S[e] - 0.5f * F[e];
}
// U (double) += D (float)
__global__ __launch_bounds__(kT) void add_f_to_d_kernel(double* U, int64_t ldu, const float* D, int64_t n) {
  const int64_t total = n * n;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < total; e += (int64_t)gridDim.x * kT) {
    const int64_t i = e % n, j = e / n;
    U[i + j * ldu] += (double)D[e];
  }
}

}  // namespace

// partial sums: one per 64 x 64 tile or per reduction block
size_t polish_parts(int64_t n) {
  const size_t t = (size_t)((n + 63) / 64);
  return std::max<size_t>(kNp, t * t);
}

size_t polish_bytes(int64_t n) {
  const size_t nn = (size_t)n * n;
  return 2 * nn * sizeof(double) + 8 * nn * sizeof(float) + (polish_parts(n) + 16) * sizeof(double) + 32 * 256;
}

int polish(const Ctx& c, double* U, int64_t ldu, const double* Xref, int64_t ldx, int64_t n, char* ws, int* passes,
           int* cg) {
  const int64_t nn = n * n;
  const hipStream_t s = c.s;
  Arena ar(ws, polish_bytes(n));
  double* G = ar.take<double>(nn);
  double* C = ar.take<double>(nn);
  float* F = ar.take<float>(nn);
  float* Hs = ar.take<float>(nn);
  float* R = ar.take<float>(nn);
  float* S = ar.take<float>(nn);
  float* P = ar.take<float>(nn);
  float* AP = ar.take<float>(nn);
  float* M = ar.take<float>(nn);
  float* U32 = ar.take<float>(nn);
  const size_t nparts = polish_parts(n);
  double* part = ar.take<double>(nparts);
  double* out = ar.take<double>(16);
  ZOLO_ARENA_OK(ar);
  const unsigned g = grid_for(nn);
  const dim3 tiles((unsigned)((n + 63) / 64), (unsigned)((n + 63) / 64));
  auto reduce_n = [&](int np, bool is_max, double* h) -> int {
    hipLaunchKernelGGL(finish_reduce_kernel, dim3(1), dim3(kT), 0, s, part, np, is_max ? 1 : 0, out);
    ZOLO_HIP(hipMemcpyAsync(h, out, sizeof(double), hipMemcpyDeviceToHost, s));
    ZOLO_HIP(hipStreamSynchronize(s));
    return kOk;
  };
  auto dot = [&](const float* a, const float* b, double* h) -> int {
    hipLaunchKernelGGL(dot_kernel, dim3(kNp), dim3(kT), 0, s, a, b, nn, part);
    return reduce_n(kNp, false, h);
  };
  // M = A_ * B_ (n x n) in the low precision of the run (FP32, or BF16x3 with
  // its ~2^-16 relative accuracy): the CG only needs S to a relative accuracy
  // well below the size of the correction it computes.
  const Prec lp = (c.low == Prec::BF16X3) ? Prec::BF16X3 : Prec::F32;
  const double rho_floor = (lp == Prec::BF16X3) ? 1e-5 : 2e-7;
  auto lgemm = [&](const float* A_, const float* B_, float* C_) -> int {
    GemmDesc d;
    d.M = n; d.N = n; d.K = n;
    d.A = A_; d.lda = n; d.B = B_; d.ldb = n; d.C = C_; d.ldc = n;
    return gemm(lp, d, s);
  };
  const int max_passes = 3;
  int pass = 0;
  for (; pass < max_passes; ++pass) {
    // G = U^T U (lower), C = U^T Xref  -- FP64 (FP64 MFMA or exact int8 products)
    GemmDesc d;
    d.transA = 1; d.M = n; d.N = n; d.K = n;
    d.A = U; d.lda = ldu; d.B = U; d.ldb = ldu; d.C = G; d.ldc = n; d.lowerC = 1;
    ZOLO_CHECK(gemm(c.big, d, s));
    d = GemmDesc();
    d.transA = 1; d.M = n; d.N = n; d.K = n;
    d.A = U; d.lda = ldu; d.B = Xref; d.ldb = ldx; d.C = C; d.ldc = n;
    d.sameA = 1;  // U is unchanged since G = U^T U: its int8 residues are reused
    ZOLO_CHECK(gemm(c.big, d, s));
    float* none = nullptr;
    hipLaunchKernelGGL(pair_tile_kernel<0>, tiles, dim3(kT), 0, s, (const double*)G, n, F, none);
    hipLaunchKernelGGL(pair_tile_kernel<1>, tiles, dim3(kT), 0, s, (const double*)C, n, Hs, R);
    // R = 2K - (F Hs - Hs F)/2 = R - 0.5 (M - M^T), M = F Hs
    ZOLO_CHECK(lgemm(F, Hs, M));
    hipLaunchKernelGGL(skew_tile_kernel<true>, tiles, dim3(kT), 0, s, R, M, n, -0.5f);
    ZOLO_HIP(hipGetLastError());
    // CG on L(S) = S Hs + Hs S = (S Hs) - (S Hs)^T for skew S
    ZOLO_HIP(hipMemsetAsync(S, 0, sizeof(float) * nn, s));
    ZOLO_HIP(hipMemcpyAsync(P, R, sizeof(float) * nn, hipMemcpyDeviceToDevice, s));
    double rr0, rr, hh;
    ZOLO_CHECK(dot(R, R, &rr0));
    ZOLO_CHECK(dot(Hs, Hs, &hh));
    rr = rr0;
    // delta = size of the correction relative to Hs.  The pass leaves an error
    // ~ delta * rho + delta^2 (rho = CG relative residual), so ask CG for no
    // more than 1e-15/delta, within what FP32 arithmetic can deliver.
    const double delta = std::sqrt(rr0 / std::max(hh, 1e-300));
    const double rho = std::max(rho_floor, std::min(1e-2, 1e-15 / std::max(delta, 1e-300)));
    const double tol2 = rho * rho * rr0;
    int it = 0;
    // per iteration: the operator product, then two passes over the n x n
    // arrays: AP = skew(P Hs) with P . AP, and the S, R update with R . R
    for (; it < 40 && rr > tol2 && rr > 0.0; ++it) {
      ZOLO_CHECK(lgemm(P, Hs, M));
      hipLaunchKernelGGL(skew_dot_kernel, tiles, dim3(kT), 0, s, AP, (const float*)M, (const float*)P, n, part);
      double pap;
      ZOLO_CHECK(reduce_n((int)(tiles.x * tiles.y), false, &pap));
      if (!(pap > 0.0)) break;
      const float a = (float)(rr / pap);
      hipLaunchKernelGGL(cg_update_kernel, dim3(kNp), dim3(kT), 0, s, S, R, (const float*)P, (const float*)AP, nn, a,
                         part);
      double rn;
      ZOLO_CHECK(reduce_n(kNp, false, &rn));
      hipLaunchKernelGGL(xpby_f_kernel, dim3(g), dim3(kT), 0, s, P, R, nn, (float)(rn / rr));
      rr = rn;
    }
    if (cg && pass < 4) cg[pass] = it;
    // E = S - F/2 (into AP) ; U += U*E
    hipLaunchKernelGGL(form_e_kernel, dim3(g), dim3(kT), 0, s, AP, S, F, nn);
    ZOLO_CHECK(k_scale_copy<double, float>(U, ldu, U32, n, n, n, 1.0, s));
    ZOLO_CHECK(lgemm(U32, AP, M));
    hipLaunchKernelGGL(add_f_to_d_kernel, dim3(g), dim3(kT), 0, s, U, ldu, M, n);
    ZOLO_HIP(hipGetLastError());
    // predicted error after this pass: delta * rho + delta^2
    if (delta * rho + delta * delta <= 2e-15) {
      ++pass;
      break;
    }
  }
  *passes = pass;
  return kOk;
}

}  // namespace zolo
