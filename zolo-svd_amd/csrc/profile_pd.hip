// Stand-alone driver of the headline polar decomposition for hardware counter
// runs (rocprofv3 --pmc): a plain HIP program over the C ABI of libzolo.so,
// with the HIP runtime libzolo.so was built against, and no Python.
//
// It first runs one copy kernel of known size (8 n^2 bytes read and written)
// so that FETCH_SIZE / WRITE_SIZE can be scaled by what they read for a
// known traffic, then builds A = Q1 diag(s) Q2^T at order n with s geometric
// from 1 to 1e-8 and the last 5 % of the values at 1e-8 (Q1, Q2 products of
// two reflections I - 2 Y (Y^T Y)^{-1} Y^T across random 256-dimensional
// subspaces, built with the library's GEMM), and runs `steps` polar
// decompositions (bounds from the first QR step's factorisation, BF16x3
// Cholesky steps, FP64 polish, H) as the benchmark does.
//
// Usage: profile_pd [n=16384] [steps=1] [precision=2]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <vector>

#include "../../include/zolo.h"

namespace {

__global__ __launch_bounds__(256) void copy_kernel(const double* __restrict__ a, double* __restrict__ b, size_t len) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < len; e += (size_t)gridDim.x * 256) b[e] = a[e];
}

__global__ __launch_bounds__(256) void fill_kernel(double* a, size_t len, unsigned long long seed) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < len; e += (size_t)gridDim.x * 256) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (e + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    a[e] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
}

__global__ __launch_bounds__(256) void eye_kernel(double* a, int64_t n) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < (size_t)n * n; e += (size_t)gridDim.x * 256)
    a[e] = (e % n == e / n) ? 1.0 : 0.0;
}

// A(:, j) *= s[j]
__global__ __launch_bounds__(256) void scale_cols_kernel(double* a, int64_t n, const double* s) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < (size_t)n * n; e += (size_t)gridDim.x * 256)
    a[e] *= s[e / n];
}

// Q <- Q (I - 2 Y M Y^T) for Y n x k, M = (Y^T Y)^{-1}: W = Q Y, Q -= 2 (W M) Y^T
int reflect(int64_t n, int64_t k, double* Q, double* Y, double* W, double* WM, double* G, unsigned long long seed) {
  fill_kernel<<<4096, 256>>>(Y, (size_t)n * k, seed);
  int st = zolo_gemm_device(0, 1, 0, k, k, n, 1.0, Y, n, Y, n, 0.0, G, k, nullptr);
  if (st) return st;
  std::vector<double> g((size_t)k * k), inv((size_t)k * k, 0.0);
  if (hipMemcpy(g.data(), G, sizeof(double) * k * k, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  for (int64_t i = 0; i < k; ++i) inv[i + i * k] = 1.0;
  for (int64_t c = 0; c < k; ++c) {  // Gauss-Jordan on the SPD Gram matrix
    const double p = g[c + c * k];
    for (int64_t j = 0; j < k; ++j) {
      g[c + j * k] /= p;
      inv[c + j * k] /= p;
    }
    for (int64_t r = 0; r < k; ++r) {
      if (r == c) continue;
      const double f = g[r + c * k];
      for (int64_t j = 0; j < k; ++j) {
        g[r + j * k] -= f * g[c + j * k];
        inv[r + j * k] -= f * inv[c + j * k];
      }
    }
  }
  if (hipMemcpy(G, inv.data(), sizeof(double) * k * k, hipMemcpyHostToDevice) != hipSuccess) return 1;
  if ((st = zolo_gemm_device(0, 0, 0, n, k, n, 1.0, Q, n, Y, n, 0.0, W, n, nullptr))) return st;
  if ((st = zolo_gemm_device(0, 0, 0, n, k, k, 1.0, W, n, G, k, 0.0, WM, n, nullptr))) return st;
  return zolo_gemm_device(0, 0, 1, n, n, k, -2.0, WM, n, Y, n, 1.0, Q, n, nullptr);
}

}  // namespace

int main(int argc, char** argv) {
  const int64_t n = argc > 1 ? std::atoll(argv[1]) : 16384;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 1;
  const int prec = argc > 3 ? std::atoi(argv[3]) : 2;
  const int64_t k = 256;
  if (n < 2 * k || n > k * k || steps < 1 || prec < 0 || prec > 4) {
    std::printf("usage: profile_pd [512 <= n <= 65536] [steps >= 1] [precision 0..4]\n");
    return 2;
  }
  if (zolo_init(0) != 0) {
    std::printf("zolo_init: %s\n", zolo_last_error());
    return 1;
  }
  const size_t nn = (size_t)n * n;
  double *A, *Q1, *Q2, *U, *H, *Y, *W, *WM, *G;
  if (hipMalloc(&A, nn * 8) || hipMalloc(&Q1, nn * 8) || hipMalloc(&Q2, nn * 8) || hipMalloc(&U, nn * 8) ||
      hipMalloc(&H, nn * 8) || hipMalloc(&Y, (size_t)n * k * 8) || hipMalloc(&W, (size_t)n * k * 8) ||
      hipMalloc(&WM, (size_t)n * k * 8) || hipMalloc(&G, (size_t)k * k * 8)) {
    std::printf("hipMalloc failed\n");
    return 1;
  }
  // calibration: one copy kernel, 8 n^2 bytes each way
  fill_kernel<<<4096, 256>>>(A, nn, 1);
  copy_kernel<<<4096, 256>>>(A, U, nn);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  std::printf("copy kernel: %zu bytes read, %zu bytes written\n", nn * 8, nn * 8);
  // Q1, Q2: products of two reflections each
  eye_kernel<<<4096, 256>>>(Q1, n);
  eye_kernel<<<4096, 256>>>(Q2, n);
  int st = 0;
  unsigned long long seed = 20260103;
  for (double* Q : {Q1, Q2})
    for (int r = 0; r < 2 && !st; ++r) st = reflect(n, k, Q, Y, W, WM, G, seed++);
  if (st) {
    std::printf("building Q: %s\n", zolo_last_error());
    return 1;
  }
  // A = Q1 diag(s) Q2^T
  const int64_t m = n - n / 20;
  std::vector<double> s(n, 1e-8);
  for (int64_t i = 0; i < m; ++i) s[i] = std::pow(10.0, -8.0 * (double)i / (double)(m - 1));
  if (hipMemcpy(G, s.data(), 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) return 1;  // n <= k^2
  scale_cols_kernel<<<4096, 256>>>(Q1, n, G);
  if ((st = zolo_gemm_device(0, 0, 1, n, n, n, 1.0, Q1, n, Q2, n, 0.0, A, n, nullptr))) {
    std::printf("forming A: %s\n", zolo_last_error());
    return 1;
  }
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  std::printf("A ready (n = %lld)\n", (long long)n);
  for (int i = 0; i < steps; ++i) {
    zolo_info_t info;
    st = zolo_pd_auto_device(prec, 'N', 0, n, A, n, U, n, H, n, &info, nullptr);
    if (st) {
      std::printf("zolo_pd_auto_device: %d %s\n", st, zolo_last_error());
      return 1;
    }
    std::printf("step %d: %.1f ms, r = %d, schedule %.*s, polish passes %d\n", i, info.time_ms, info.r, info.iters,
                info.schedule, info.polish_passes);
  }
  zolo_finalize();
  std::printf("profile_pd done\n");
  return 0;
}
