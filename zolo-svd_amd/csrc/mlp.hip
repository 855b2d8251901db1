// Learned lower bound of sigma_min for the batched path: matrix features, a
// 16 -> 64 -> 64 -> 1 tanh network, and its training (pinball loss, Adam).
//
// Features of X = A / ||A||_F (one workgroup per matrix, FP64 arithmetic on the
// FP32 data, deterministic start and sketch vectors):
//   f0      log2 n
//   f1..f8  log10 of the 8 Ritz values of X^T X after 8 Lanczos steps (full
//           reorthogonalisation), descending
//   f9      log10 ||X||_1            f10  log10 ||X||_inf
//   f11..14 log10 |r_kk| of the R factor of X Omega (Omega n x 4 Gaussian sketch)
//   f15     n theta_1 = n ||X||_2^2 / ||X||_F^2 estimate (n / stable rank)
// and smax = sqrt(theta_1), the Lanczos estimate of sigma_max / ||A||_F (from
// below).  The network predicts y = log10(sigma_min / sigma_max); the bound is
// beta = 10^(y - margin) smax.
//
// Parameter vector (float): mu[16] sd[16] | W1[64][16] b1[64] W2[64][64] b2[64]
// w3[64] b3 | margin.  Only the middle block is trained.
#include <cmath>

#include "batched.hpp"
#include "mlp.hpp"

namespace zolo {
namespace {

constexpr int kT = 256;
constexpr int kLanczos = 8;
constexpr int kSketch = 4;

__device__ inline double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

__device__ inline double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
  __syncthreads();
  return r;
}

// Deterministic pseudo-random numbers (a 32-bit integer hash).
__device__ inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
__device__ inline double unit_uniform(uint32_t key) { return ((double)hash32(key) + 0.5) * (1.0 / 4294967296.0); }
__device__ inline double gauss(uint32_t key) {
  const double u1 = unit_uniform(2 * key + 1), u2 = unit_uniform(2 * key + 2);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Eigenvalue e (descending) of the k x k symmetric tridiagonal (a, b) by
// bisection on the Sturm count.
__device__ double tridiag_eig(const double* a, const double* b, int k, int e) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i + 1 < k ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  const int want = k - e;  // number of eigenvalues < x must reach `want` at the (k-e)-th smallest
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    int cnt = 0;
    double d = 1.0;
    for (int i = 0; i < k; ++i) {
      d = a[i] - mid - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.0);
      if (d == 0.0) d = 1e-300;
      if (d < 0.0) ++cnt;
    }
    if (cnt >= want) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Sum over j < n of f(A[i + j ld], j) for one row i, eight loads in flight
// (the loads of a group are issued before any is used) and eight partial
// sums, added in a fixed order.
template <typename F>
__device__ inline double row_sum8(const float* A, int64_t ld, int i, int n, F&& f) {
  double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const float* a = A + i;
  int j = 0;
  for (; j + 8 <= n; j += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = a[(int64_t)(j + u) * ld];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] += f((double)v[u], j + u);
  }
  for (; j < n; ++j) s[j & 7] += f((double)a[(int64_t)j * ld], j);
  return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}
// Sum over i < n (1 <= n <= kMaxBatchN) of f(A[i + j ld], i) for one
// column j, lanes over rows: the kMaxBatchN / 64 loads of a lane all in
// flight (rows past n read row n - 1 and are left out of the sum).
template <typename F>
__device__ inline double col_sum_wave(const float* A, int64_t ld, int j, int n, int lane, F&& f) {
  constexpr int R = kMaxBatchN / 64;
  const float* a = A + (int64_t)j * ld;
  float v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = a[min(lane + 64 * r, n - 1)];
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (lane + 64 * r < n) s += f((double)v[r], lane + 64 * r);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

__global__ __launch_bounds__(kT) void features_kernel(const BatchMat* mats, const double* alpha, float* feat,
                                                      float* smax) {
  __shared__ double Q[kLanczos + 1][kMaxBatchN];
  __shared__ double u[kMaxBatchN];
  __shared__ double w[kMaxBatchN];
  __shared__ double Y[kSketch][kMaxBatchN];
  __shared__ double Om[kSketch][kMaxBatchN];
  __shared__ double ta[kLanczos], tb[kLanczos], theta[kLanczos];
  __shared__ double sh[4];
  const int g = blockIdx.x;
  const BatchMat m = mats[g];
  const int n = (int)m.n;
  const int t = threadIdx.x;
  const double ia = 1.0 / alpha[g];
  const float* A = m.A;
  const int64_t ld = m.lda;
  auto x = [&](int i, int j) { return ia * (double)A[i + (int64_t)j * ld]; };

  // norms: row sums (thread per row), column sums (one wave per column,
  // lanes over rows: coalesced)
  double rmax = 0.0, cmax = 0.0;
  for (int i = t; i < n; i += kT)
    rmax = fmax(rmax, row_sum8(A, ld, i, n, [&](double a, int) { return fabs(ia * a); }));
  const int lane = t & 63, wave = t >> 6;
  for (int j = wave; j < n; j += kT / 64)
    cmax = fmax(cmax, col_sum_wave(A, ld, j, n, lane, [&](double a, int) { return fabs(ia * a); }));
  const double ninf = block_max(rmax, sh);
  const double n1 = block_max(cmax, sh);

  // sketch Y = X Omega and the diagonal of its R factor (modified Gram-Schmidt);
  // Omega is generated once per matrix into LDS
  for (int j = t; j < n; j += kT)
#pragma unroll
    for (int k = 0; k < kSketch; ++k) Om[k][j] = gauss((uint32_t)(j * kSketch + k));
  __syncthreads();
  for (int i = t; i < n; i += kT) {
    double acc[kSketch] = {0.0, 0.0, 0.0, 0.0};
    const float* a = A + i;
    int j = 0;
    for (; j + 8 <= n; j += 8) {  // eight loads in flight
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = a[(int64_t)(j + u) * ld];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double xv = ia * (double)v[u];
#pragma unroll
        for (int k = 0; k < kSketch; ++k) acc[k] += xv * Om[k][j + u];
      }
    }
    for (; j < n; ++j) {
      const double xv = x(i, j);
#pragma unroll
      for (int k = 0; k < kSketch; ++k) acc[k] += xv * Om[k][j];
    }
#pragma unroll
    for (int k = 0; k < kSketch; ++k) Y[k][i] = acc[k];
  }
  __syncthreads();
  double rdiag[kSketch];
  for (int k = 0; k < kSketch; ++k) {
    for (int p = 0; p < k; ++p) {
      double d = 0.0;
      for (int i = t; i < n; i += kT) d += Y[p][i] * Y[k][i];
      d = block_sum(d, sh);
      for (int i = t; i < n; i += kT) Y[k][i] -= d * Y[p][i];
      __syncthreads();
    }
    double s = 0.0;
    for (int i = t; i < n; i += kT) s += Y[k][i] * Y[k][i];
    s = sqrt(block_sum(s, sh));
    rdiag[k] = s;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    for (int i = t; i < n; i += kT) Y[k][i] *= inv;
    __syncthreads();
  }

  // Lanczos on X^T X with full reorthogonalisation
  for (int i = t; i < n; i += kT) Q[0][i] = ((hash32((uint32_t)i + 0x9e3779b9U) & 1U) ? 1.0 : -1.0) / sqrt((double)n);
  __syncthreads();
  int k = 0;
  for (; k < kLanczos && k < n; ++k) {
    for (int i = t; i < n; i += kT)  // u = X q_k
      u[i] = row_sum8(A, ld, i, n, [&](double a, int j) { return ia * a * Q[k][j]; });
    __syncthreads();
    for (int j = wave; j < n; j += kT / 64) {  // w = X^T u, one wave per column
      const double s = col_sum_wave(A, ld, j, n, lane, [&](double a, int i) { return ia * a * u[i]; });
      if (lane == 0) w[j] = s;
    }
    __syncthreads();
    double a = 0.0;
    for (int i = t; i < n; i += kT) a += Q[k][i] * w[i];
    a = block_sum(a, sh);
    if (t == 0) ta[k] = a;
    for (int p = 0; p <= k; ++p) {  // w -= Q_p (Q_p^T w), twice for stability
      for (int pass = 0; pass < 2; ++pass) {
        double d = 0.0;
        for (int i = t; i < n; i += kT) d += Q[p][i] * w[i];
        d = block_sum(d, sh);
        for (int i = t; i < n; i += kT) w[i] -= d * Q[p][i];
        __syncthreads();
      }
    }
    double bnorm = 0.0;
    for (int i = t; i < n; i += kT) bnorm += w[i] * w[i];
    bnorm = sqrt(block_sum(bnorm, sh));
    if (t == 0) tb[k] = bnorm;
    __syncthreads();
    if (bnorm <= 1e-14 * fabs(a) || k + 1 == kLanczos) {
      ++k;
      break;
    }
    for (int i = t; i < n; i += kT) Q[k + 1][i] = w[i] / bnorm;
    __syncthreads();
  }
  const int kk = k;  // Lanczos steps done (<= kLanczos)
  if (t < kk) theta[t] = tridiag_eig(ta, tb, kk, t);
  __syncthreads();
  if (t == 0) {
    float* f = feat + (int64_t)g * kFeat;
    const double th1 = fmax(theta[0], 1e-300);
    f[0] = (float)log2((double)n);
    for (int e = 0; e < kLanczos; ++e) {
      const double th = (e < kk) ? theta[e] : theta[kk - 1];
      f[1 + e] = (float)log10(fmax(th, 1e-30));
    }
    f[9] = (float)log10(fmax(n1, 1e-30));
    f[10] = (float)log10(fmax(ninf, 1e-30));
    for (int e = 0; e < kSketch; ++e) f[11 + e] = (float)log10(fmax(rdiag[e], 1e-30));
    f[15] = (float)(n * th1);
    smax[g] = (float)sqrt(th1);
  }
}

// ---------------------------------------------------------------------------
// Network: forward, loss, backward, parameter gradients, Adam.
// Four samples per workgroup, one hidden unit per thread.
// ---------------------------------------------------------------------------
struct Net {
  const float* mu;
  const float* sd;
  const float* W1;
  const float* b1;
  const float* W2;
  const float* b2;
  const float* w3;
  const float* b3;
};
__device__ inline Net net_of(const float* P) {
  Net q;
  q.mu = P;
  q.sd = P + kFeat;
  q.W1 = P + kParamNet0;
  q.b1 = q.W1 + kHid * kFeat;
  q.W2 = q.b1 + kHid;
  q.b2 = q.W2 + kHid * kHid;
  q.w3 = q.b2 + kHid;
  q.b3 = q.w3 + kHid;
  return q;
}

__global__ __launch_bounds__(kT) void mlp_forward_kernel(const float* P, const float* X, int N, float* H1, float* H2,
                                                         float* Y) {
  __shared__ float xs[4][kFeat];
  __shared__ float h1s[4][kHid];
  const Net q = net_of(P);
  const int s = threadIdx.x >> 6, u = threadIdx.x & 63;
  const int n = blockIdx.x * 4 + s;
  const bool live = n < N;
  if (u < kFeat) xs[s][u] = live ? (X[(int64_t)n * kFeat + u] - q.mu[u]) / q.sd[u] : 0.0f;
  __syncthreads();
  float z = q.b1[u];
#pragma unroll
  for (int f = 0; f < kFeat; ++f) z += q.W1[u * kFeat + f] * xs[s][f];
  const float h1 = tanhf(z);
  h1s[s][u] = h1;
  __syncthreads();
  float z2 = q.b2[u];
  for (int j = 0; j < kHid; ++j) z2 += q.W2[u * kHid + j] * h1s[s][j];
  const float h2 = tanhf(z2);
  float y = q.w3[u] * h2;
  for (int o = 32; o > 0; o >>= 1) y += __shfl_down(y, o, 64);
  if (live) {
    if (H1) H1[(int64_t)n * kHid + u] = h1;
    if (H2) H2[(int64_t)n * kHid + u] = h2;
    if (u == 0) Y[n] = y + q.b3[0];
  }
}

// dL/dy of the pinball loss mean_n rho_tau(t_n - y_n), then the hidden deltas.
__global__ __launch_bounds__(kT) void mlp_backward_kernel(const float* P, const float* T, int N, float tau,
                                                          const float* H1, const float* H2, const float* Y, float* D3,
                                                          float* D2, float* D1) {
  __shared__ float d2s[4][kHid];
  const Net q = net_of(P);
  const int s = threadIdx.x >> 6, u = threadIdx.x & 63;
  const int n = blockIdx.x * 4 + s;
  const bool live = n < N;
  float d3 = 0.0f;
  if (live) d3 = ((Y[n] > T[n] ? 1.0f : 0.0f) - tau) / (float)N;
  const float h2 = live ? H2[(int64_t)n * kHid + u] : 0.0f;
  const float d2 = d3 * q.w3[u] * (1.0f - h2 * h2);
  d2s[s][u] = d2;
  __syncthreads();
  float g = 0.0f;
  for (int k = 0; k < kHid; ++k) g += q.W2[k * kHid + u] * d2s[s][k];
  const float h1 = live ? H1[(int64_t)n * kHid + u] : 0.0f;
  const float d1 = g * (1.0f - h1 * h1);
  if (live) {
    D2[(int64_t)n * kHid + u] = d2;
    D1[(int64_t)n * kHid + u] = d1;
    if (u == 0) D3[n] = d3;
  }
}

// One thread per trained parameter: sum over samples in a fixed order.
__global__ __launch_bounds__(kT) void mlp_param_grad_kernel(const float* P, const float* X, int N, const float* H1,
                                                            const float* H2, const float* D3, const float* D2,
                                                            const float* D1, float* G) {
  const int p = blockIdx.x * kT + threadIdx.x;
  if (p >= kParamTrain) return;
  const Net q = net_of(P);
  double acc = 0.0;
  if (p < kHid * kFeat) {  // W1[u][f]
    const int u = p / kFeat, f = p % kFeat;
    const float mu = q.mu[f], isd = 1.0f / q.sd[f];
    for (int n = 0; n < N; ++n) acc += (double)D1[(int64_t)n * kHid + u] * (double)((X[(int64_t)n * kFeat + f] - mu) * isd);
  } else if (p < kHid * kFeat + kHid) {  // b1
    const int u = p - kHid * kFeat;
    for (int n = 0; n < N; ++n) acc += D1[(int64_t)n * kHid + u];
  } else if (p < kHid * kFeat + kHid + kHid * kHid) {  // W2[u][j]
    const int r = p - kHid * kFeat - kHid;
    const int u = r / kHid, j = r % kHid;
    for (int n = 0; n < N; ++n) acc += (double)D2[(int64_t)n * kHid + u] * (double)H1[(int64_t)n * kHid + j];
  } else if (p < kHid * kFeat + 2 * kHid + kHid * kHid) {  // b2
    const int u = p - kHid * kFeat - kHid - kHid * kHid;
    for (int n = 0; n < N; ++n) acc += D2[(int64_t)n * kHid + u];
  } else if (p < kHid * kFeat + 3 * kHid + kHid * kHid) {  // w3
    const int u = p - kHid * kFeat - 2 * kHid - kHid * kHid;
    for (int n = 0; n < N; ++n) acc += (double)D3[n] * (double)H2[(int64_t)n * kHid + u];
  } else {  // b3
    for (int n = 0; n < N; ++n) acc += D3[n];
  }
  G[p] = (float)acc;
}

__global__ __launch_bounds__(kT) void pinball_loss_kernel(const float* T, const float* Y, int N, float tau,
                                                          float* out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int n = threadIdx.x; n < N; n += kT) {
    const double r = (double)T[n] - (double)Y[n];
    s += r >= 0.0 ? tau * r : (tau - 1.0) * r;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *out = (float)(s / N);
}

__global__ __launch_bounds__(kT) void adam_kernel(float* P, const float* G, float* M, float* V, float lr, float b1,
                                                  float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * kT + threadIdx.x;
  if (p >= kParamTrain) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.0f - b1) * g;
  const float v = b2 * V[p] + (1.0f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[kParamNet0 + p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// beta_g = 10^(y_g - margin) smax_g
__global__ __launch_bounds__(kT) void beta_kernel(const float* P, const float* Y, const float* smax, int N,
                                                  double* beta) {
  const int n = blockIdx.x * kT + threadIdx.x;
  if (n >= N) return;
  const double y = (double)Y[n] - (double)P[kParamMargin];
  beta[n] = pow(10.0, fmin(y, 0.0)) * (double)smax[n];
}

}  // namespace

int kl_features(const BatchMat* mats, const double* alpha, int count, float* feat, float* smax, hipStream_t s) {
  hipLaunchKernelGGL(features_kernel, dim3(count), dim3(kT), 0, s, mats, alpha, feat, smax);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kl_forward(const float* P, const float* X, int N, float* H1, float* H2, float* Y, hipStream_t s) {
  if (N <= 0) return kOk;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((N + 3) / 4), dim3(kT), 0, s, P, X, N, H1, H2, Y);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kl_grad(const float* P, const float* X, const float* T, int N, float tau, float* H1, float* H2, float* Y,
            float* D3, float* D2, float* D1, float* G, float* loss, hipStream_t s) {
  ZOLO_CHECK(kl_forward(P, X, N, H1, H2, Y, s));
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((N + 3) / 4), dim3(kT), 0, s, P, T, N, tau, H1, H2, Y, D3, D2, D1);
  hipLaunchKernelGGL(mlp_param_grad_kernel, dim3((kParamTrain + kT - 1) / kT), dim3(kT), 0, s, P, X, N, H1, H2, D3, D2,
                     D1, G);
  if (loss) hipLaunchKernelGGL(pinball_loss_kernel, dim3(1), dim3(kT), 0, s, T, Y, N, tau, loss);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kl_adam(float* P, const float* G, float* M, float* V, float lr, int step, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((kParamTrain + kT - 1) / kT), dim3(kT), 0, s, P, G, M, V, lr, b1, b2, eps, c1,
                     c2);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

int kl_beta(const float* P, const float* Y, const float* smax, int N, double* beta, hipStream_t s) {
  hipLaunchKernelGGL(beta_kernel, dim3((N + kT - 1) / kT), dim3(kT), 0, s, P, Y, smax, N, beta);
  ZOLO_HIP(hipGetLastError());
  return kOk;
}

}  // namespace zolo
