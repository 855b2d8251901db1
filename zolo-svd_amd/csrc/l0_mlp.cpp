// Learned lower bound of sigma_min for the batched path (host side): the
// network parameters live on the device once set; the bound of a batch is
// features -> forward -> beta in three launches.
#include "mlp.hpp"

#include <chrono>
#include <vector>

namespace zolo {

namespace {
float* g_params = nullptr;  // device copy of the parameter vector (kParamCount floats)
bool g_have_params = false;
}  // namespace

int l0_set_weights(const float* host_params) {
  if (!g_params) ZOLO_HIP(hipMalloc(&g_params, sizeof(float) * kParamCount));
  ZOLO_HIP(hipMemcpy(g_params, host_params, sizeof(float) * kParamCount, hipMemcpyHostToDevice));
  g_have_params = true;
  return kOk;
}

void l0_release() {
  if (g_params) (void)hipFree(g_params);
  g_params = nullptr;
  g_have_params = false;
}

int l0_features(int count, const int64_t* n, const float* const* A, const int64_t* lda, float* feat, float* smax,
                hipStream_t s) {
  const size_t bytes = 4 * 256 + sizeof(BatchMat) * (size_t)count + sizeof(double) * (size_t)count;
  char* ws = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &ws));
  Arena ar(ws, bytes);
  BatchMat* d_mats = ar.take<BatchMat>(count);
  double* d_alpha = ar.take<double>(count);
  ZOLO_ARENA_OK(ar);
  std::vector<BatchMat> h(count);
  for (int i = 0; i < count; ++i) h[i] = BatchMat{n[i], A[i], lda[i], nullptr, 0};
  ZOLO_HIP(hipMemcpyAsync(d_mats, h.data(), sizeof(BatchMat) * count, hipMemcpyHostToDevice, s));
  ZOLO_CHECK(kb_alpha(d_mats, count, d_alpha, s));
  ZOLO_CHECK(kl_features(d_mats, d_alpha, count, feat, smax, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  return kOk;
}

int l0_learned_beta(int count, const int64_t* n, const float* const* A, const int64_t* lda,
                    std::vector<double>& beta, double* ms, hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!g_have_params) {
    set_error("learned l0: no network weights set (zolo_l0_set_weights)");
    return kErrArg;
  }
  beta.assign(count, 0.0);
  if (count == 0) return kOk;
  const size_t bytes = 8 * 256 + sizeof(BatchMat) * (size_t)count +
                       (size_t)count * (2 * sizeof(double) + (kFeat + 2) * sizeof(float));
  char* ws = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &ws));
  Arena ar(ws, bytes);
  BatchMat* d_mats = ar.take<BatchMat>(count);
  double* d_alpha = ar.take<double>(count);
  double* d_beta = ar.take<double>(count);
  float* feat = ar.take<float>((size_t)count * kFeat);
  float* smax = ar.take<float>(count);
  float* y = ar.take<float>(count);
  ZOLO_ARENA_OK(ar);
  std::vector<BatchMat> h(count);
  for (int i = 0; i < count; ++i) h[i] = BatchMat{n[i], A[i], lda[i], nullptr, 0};
  ZOLO_HIP(hipMemcpyAsync(d_mats, h.data(), sizeof(BatchMat) * count, hipMemcpyHostToDevice, s));
  ZOLO_CHECK(kb_alpha(d_mats, count, d_alpha, s));
  ZOLO_CHECK(kl_features(d_mats, d_alpha, count, feat, smax, s));
  ZOLO_CHECK(kl_forward(g_params, feat, count, nullptr, nullptr, y, s));
  ZOLO_CHECK(kl_beta(g_params, y, smax, count, d_beta, s));
  ZOLO_HIP(hipMemcpyAsync(beta.data(), d_beta, sizeof(double) * count, hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return kOk;
}

namespace {
struct TrainScratch {
  float *H1, *H2, *Y, *D3, *D2, *D1, *G, *M, *V, *loss;
};
int take_train(int N, TrainScratch& t) {
  const size_t bytes = 12 * 256 + sizeof(float) * ((size_t)N * (4 * kHid + 2) + 3 * (size_t)kParamTrain + 1);
  char* ws = nullptr;
  ZOLO_CHECK(workspace_acquire(bytes, &ws));
  Arena ar(ws, bytes);
  t.H1 = ar.take<float>((size_t)N * kHid);
  t.H2 = ar.take<float>((size_t)N * kHid);
  t.D2 = ar.take<float>((size_t)N * kHid);
  t.D1 = ar.take<float>((size_t)N * kHid);
  t.Y = ar.take<float>(N);
  t.D3 = ar.take<float>(N);
  t.G = ar.take<float>(kParamTrain);
  t.M = ar.take<float>(kParamTrain);
  t.V = ar.take<float>(kParamTrain);
  t.loss = ar.take<float>(1);
  ZOLO_ARENA_OK(ar);
  return kOk;
}
}  // namespace

int l0_grad(const float* P, const float* X, const float* T, int N, float tau, float* grad, float* loss,
            hipStream_t s) {
  TrainScratch t;
  ZOLO_CHECK(take_train(N, t));
  ZOLO_CHECK(kl_grad(P, X, T, N, tau, t.H1, t.H2, t.Y, t.D3, t.D2, t.D1, grad, t.loss, s));
  if (loss) ZOLO_HIP(hipMemcpyAsync(loss, t.loss, sizeof(float), hipMemcpyDeviceToHost, s));
  ZOLO_HIP(hipStreamSynchronize(s));
  return kOk;
}

int l0_train(float* P, const float* X, const float* T, int N, int steps, float lr, float tau, float* loss_hist,
             hipStream_t s) {
  TrainScratch t;
  ZOLO_CHECK(take_train(N, t));
  ZOLO_HIP(hipMemsetAsync(t.M, 0, sizeof(float) * kParamTrain, s));
  ZOLO_HIP(hipMemsetAsync(t.V, 0, sizeof(float) * kParamTrain, s));
  for (int k = 1; k <= steps; ++k) {
    ZOLO_CHECK(kl_grad(P, X, T, N, tau, t.H1, t.H2, t.Y, t.D3, t.D2, t.D1, t.G, loss_hist ? t.loss : nullptr, s));
    if (loss_hist) ZOLO_HIP(hipMemcpyAsync(loss_hist + (k - 1), t.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    ZOLO_CHECK(kl_adam(P, t.G, t.M, t.V, lr, k, s));
  }
  ZOLO_HIP(hipStreamSynchronize(s));
  return kOk;
}

}  // namespace zolo
