// Learned lower bound: features, network and training kernels (mlp.hip).
#pragma once
#include "batched.hpp"

namespace zolo {

constexpr int kFeat = 16;                                  // input features
constexpr int kHid = 64;                                   // hidden width (two layers)
constexpr int kParamNet0 = 2 * kFeat;                      // after mu[16], sd[16]
constexpr int kParamTrain = kHid * kFeat + kHid + kHid * kHid + kHid + kHid + 1;  // W1 b1 W2 b2 w3 b3
constexpr int kParamMargin = kParamNet0 + kParamTrain;     // margin (decades)
constexpr int kParamCount = kParamMargin + 1;

int kl_features(const BatchMat* mats, const double* alpha, int count, float* feat, float* smax, hipStream_t s);
int kl_forward(const float* P, const float* X, int N, float* H1, float* H2, float* Y, hipStream_t s);
int kl_grad(const float* P, const float* X, const float* T, int N, float tau, float* H1, float* H2, float* Y,
            float* D3, float* D2, float* D1, float* G, float* loss, hipStream_t s);
int kl_adam(float* P, const float* G, float* M, float* V, float lr, int step, hipStream_t s);
int kl_beta(const float* P, const float* Y, const float* smax, int N, double* beta, hipStream_t s);

// Host side (l0_mlp.cpp).
int l0_set_weights(const float* host_params);
void l0_release();
int l0_features(int count, const int64_t* n, const float* const* A, const int64_t* lda, float* feat, float* smax,
                hipStream_t s);
int l0_grad(const float* P, const float* X, const float* T, int N, float tau, float* grad, float* loss, hipStream_t s);
int l0_train(float* P, const float* X, const float* T, int N, int steps, float lr, float tau, float* loss_hist,
             hipStream_t s);

}  // namespace zolo
