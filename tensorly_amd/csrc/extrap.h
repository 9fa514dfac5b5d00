// Learned ALS extrapolation: the 16 input features of the error history and
// the evaluation of the 16 -> 64 -> 64 -> 1 tanh network that maps them to
// the step alpha of A <- A + alpha (A - A_prev).  Shared by the single-problem
// driver (cp_als.hip) and the batched persistent kernel (batched.hip), so the
// step the network was trained on is exactly the step it is used for.
#pragma once
#include "common.h"

namespace tla {

constexpr int kMlpIn = 16, kMlpH = 64;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;

// History arrays are indexed by sweep slot: errors[s] (relative error after
// slot s), alphas[s] (step applied before slot s), accepted[s] (0 if slot s
// was an extrapolated sweep that got reverted).
__device__ __noinline__ void extrap_features(const double* errors, const float* alphas, const int* accepted, int s,
                                       float* f) {
  double e[5];
  for (int k = 0; k < 5; ++k) e[k] = errors[max(s - k, 0)];
  double d[4];
  for (int k = 0; k < 4; ++k) d[k] = e[k + 1] - e[k];
  const double e0 = fmax(e[0], 1e-30);
  f[0] = (float)log10(e0);
  for (int k = 0; k < 4; ++k) f[1 + k] = (float)log10(fmax(fabs(d[k]), 1e-30) / e0);
  for (int k = 0; k < 3; ++k) f[5 + k] = (d[k + 1] != 0.0) ? (float)fmin(fmax(d[k] / d[k + 1], -2.0), 2.0) : 0.0f;
  const float a1 = (s >= 1) ? alphas[s - 1] : 0.0f;
  const float a2 = (s >= 3) ? alphas[s - 3] : 0.0f;
  f[8] = tanhf(a1 / 4.0f);
  f[9] = tanhf(a2 / 4.0f);
  f[10] = (s >= 1) ? (float)accepted[s - 1] : 0.5f;
  float acc = 0.0f;
  int cnt = 0;
  for (int k = 1; k <= 7; k += 2)
    if (s - k >= 0) {
      acc += (float)accepted[s - k];
      ++cnt;
    }
  f[11] = cnt ? acc / cnt : 0.5f;
  f[12] = (float)(log10((double)s + 1.0) / 3.0);
  f[13] = (d[0] > 0.0) ? 1.0f : ((d[0] < 0.0) ? -1.0f : 0.0f);
  f[14] = (float)(fmin(fmax((e[4] - e[0]) / (4.0 * fmax(fabs(d[0]), 1e-30)), -4.0), 4.0) / 4.0);
  f[15] = 1.0f;
}

// Network output y for the features f[16] (LDS), evaluated by threads
// t < 64 of the calling block; h1/part are 64-float LDS scratch.  Every
// thread of the block must call it (it synchronises); the result is valid in
// thread 0 only.  alpha = clip(2^y - 1, 0, alpha_max), NaN -> 0.
__device__ __noinline__ float mlp_alpha_block(const float* __restrict__ params, const float* f, float* h1, float* part,
                                        float alpha_max) {
  const int t = threadIdx.x;
  const float* W1 = params;
  const float* b1 = W1 + kMlpH * kMlpIn;
  const float* W2 = b1 + kMlpH;
  const float* b2 = W2 + kMlpH * kMlpH;
  const float* w3 = b2 + kMlpH;
  const float* b3 = w3 + kMlpH;
  if (t < kMlpH) {
    float z = b1[t];
    for (int i = 0; i < kMlpIn; ++i) z = fmaf(W1[t * kMlpIn + i], f[i], z);
    h1[t] = tanhf(z);
  }
  __syncthreads();
  if (t < kMlpH) {
    float z = b2[t];
    for (int j = 0; j < kMlpH; ++j) z = fmaf(W2[t * kMlpH + j], h1[j], z);
    part[t] = w3[t] * tanhf(z);
  }
  __syncthreads();
  float a = 0.0f;
  if (t == 0) {
    float y = b3[0];
    for (int k = 0; k < kMlpH; ++k) y += part[k];
    a = exp2f(y) - 1.0f;
    a = fminf(fmaxf(a, 0.0f), alpha_max);
    if (!(a == a)) a = 0.0f;
  }
  return a;
}

}  // namespace tla
