// The learned ALS extrapolation network: 16 -> 64 -> 64 -> 1 MLP with tanh
// hidden activations and a linear output, trained by mean-squared error with
// Adam.  Parameter vector layout (5313 floats):
//   W1 [64][16] | b1 [64] | W2 [64][64] | b2 [64] | w3 [64] | b3 [1]
//
// mlp_forward  : one thread per sample, parameters in LDS.
// mlp_train    : one persistent workgroup runs `steps` minibatch steps over
//                the dataset (rows taken in the order given by `order`,
//                wrapping around), all activations of a minibatch in LDS,
//                gradients reduced over the batch in a fixed order (bitwise
//                reproducible), Adam update in place.
#include <math.h>

#include "common.h"

namespace tla {

constexpr int kIn = 16, kH = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kNParams = kOffB3 + 1;  // 5313
constexpr int kMaxBatch = 128;
constexpr int kTrainThreads = 256;

__device__ __forceinline__ float mlp_sample(const float* P, const float* x) {
  float h1[kH];
#pragma unroll 4
  for (int j = 0; j < kH; ++j) {
    float z = P[kOffB1 + j];
#pragma unroll
    for (int f = 0; f < kIn; ++f) z = fmaf(P[kOffW1 + j * kIn + f], x[f], z);
    h1[j] = tanhf(z);
  }
  float y = P[kOffB3];
  for (int k = 0; k < kH; ++k) {
    float z = P[kOffB2 + k];
#pragma unroll 8
    for (int j = 0; j < kH; ++j) z = fmaf(P[kOffW2 + k * kH + j], h1[j], z);
    y = fmaf(P[kOffW3 + k], tanhf(z), y);
  }
  return y;
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* __restrict__ params, const float* __restrict__ X,
                                                          int n, float* __restrict__ out) {
  __shared__ float P[kNParams];
  for (int i = threadIdx.x; i < kNParams; i += blockDim.x) P[i] = params[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn];
#pragma unroll
  for (int f = 0; f < kIn; ++f) x[f] = X[(int64_t)s * kIn + f];
  out[s] = mlp_sample(P, x);
}

struct TrainArgs {
  float* params;
  float* adam_m;
  float* adam_v;
  const float* X;
  const float* y;
  const int* order;  // dataset row order (length n)
  int n, batch, steps;
  float lr, beta1, beta2, eps;
  int t0;            // Adam step counter before this call
  float* loss;       // [steps] mean squared error of each minibatch (before the update)
  float* grad_out;   // if set: write the first minibatch's gradient here and stop (no update)
};

__global__ __launch_bounds__(kTrainThreads) void mlp_train_kernel(TrainArgs a) {
  extern __shared__ __align__(16) float sm[];
  float* P = sm;                          // kNParams
  float* G = P + kNParams;                // kNParams gradient
  float* Xb = G + kNParams;               // batch x 16
  float* H1 = Xb + kMaxBatch * kIn;       // batch x 64  (then delta1)
  float* H2 = H1 + kMaxBatch * kH;        // batch x 64  (then delta2)
  float* gy = H2 + kMaxBatch * kH;        // batch
  float* red = gy + kMaxBatch;            // 256
  const int tid = threadIdx.x;
  const int B = a.batch;
  for (int i = tid; i < kNParams; i += kTrainThreads) P[i] = a.params[i];
  __syncthreads();
  for (int step = 0; step < a.steps; ++step) {
    // gather the minibatch
    for (int e = tid; e < B * kIn; e += kTrainThreads) {
      const int i = e / kIn, f = e % kIn;
      const int row = a.order[((int64_t)step * B + i) % a.n];
      Xb[e] = a.X[(int64_t)row * kIn + f];
    }
    __syncthreads();
    // forward: H1 = tanh(X W1^T + b1)
    for (int e = tid; e < B * kH; e += kTrainThreads) {
      const int i = e / kH, j = e % kH;
      float z = P[kOffB1 + j];
#pragma unroll
      for (int f = 0; f < kIn; ++f) z = fmaf(P[kOffW1 + j * kIn + f], Xb[i * kIn + f], z);
      H1[e] = tanhf(z);
    }
    __syncthreads();
    // H2 = tanh(H1 W2^T + b2)
    for (int e = tid; e < B * kH; e += kTrainThreads) {
      const int i = e / kH, k = e % kH;
      float z = P[kOffB2 + k];
      for (int j = 0; j < kH; ++j) z = fmaf(P[kOffW2 + k * kH + j], H1[i * kH + j], z);
      H2[e] = tanhf(z);
    }
    __syncthreads();
    // output and loss gradient dL/dy = 2 (y_hat - y) / B
    if (tid < B) {
      float yh = P[kOffB3];
      for (int k = 0; k < kH; ++k) yh = fmaf(P[kOffW3 + k], H2[tid * kH + k], yh);
      const int row = a.order[((int64_t)step * B + tid) % a.n];
      const float diff = yh - a.y[row];
      gy[tid] = 2.0f * diff / (float)B;
      red[tid] = diff * diff;
    }
    __syncthreads();
    if (tid == 0) {
      float s = 0.0f;
      for (int i = 0; i < B; ++i) s += red[i];
      if (a.loss) a.loss[step] = s / (float)B;
    }
    // dw3, db3
    if (tid < kH) {
      float s = 0.0f;
      for (int i = 0; i < B; ++i) s = fmaf(gy[i], H2[i * kH + tid], s);
      G[kOffW3 + tid] = s;
    } else if (tid == kH) {
      float s = 0.0f;
      for (int i = 0; i < B; ++i) s += gy[i];
      G[kOffB3] = s;
    }
    __syncthreads();
    // delta2 = gy w3 (1 - H2^2), in place of H2
    for (int e = tid; e < B * kH; e += kTrainThreads) {
      const int i = e / kH, k = e % kH;
      const float h = H2[e];
      H2[e] = gy[i] * P[kOffW3 + k] * (1.0f - h * h);
    }
    __syncthreads();
    // dW2[k][j] = sum_i delta2[i][k] H1[i][j]; db2[k] = sum_i delta2[i][k]
    for (int e = tid; e < kH * kH; e += kTrainThreads) {
      const int k = e / kH, j = e % kH;
      float s = 0.0f;
      for (int i = 0; i < B; ++i) s = fmaf(H2[i * kH + k], H1[i * kH + j], s);
      G[kOffW2 + e] = s;
    }
    if (tid < kH) {
      float s = 0.0f;
      for (int i = 0; i < B; ++i) s += H2[i * kH + tid];
      G[kOffB2 + tid] = s;
    }
    __syncthreads();
    // delta1 = (delta2 W2) (1 - H1^2): computed into registers first, then
    // written over H1 once every thread has read H1
    float d1[(kMaxBatch * kH + kTrainThreads - 1) / kTrainThreads];
    int cnt = 0;
    for (int e = tid; e < B * kH; e += kTrainThreads, ++cnt) {
      const int i = e / kH, j = e % kH;
      float s = 0.0f;
      for (int k = 0; k < kH; ++k) s = fmaf(H2[i * kH + k], P[kOffW2 + k * kH + j], s);
      const float h = H1[e];
      d1[cnt] = s * (1.0f - h * h);
    }
    __syncthreads();
    cnt = 0;
    for (int e = tid; e < B * kH; e += kTrainThreads, ++cnt) H1[e] = d1[cnt];
    __syncthreads();
    // dW1[j][f] = sum_i delta1[i][j] X[i][f]; db1[j] = sum_i delta1[i][j]
    for (int e = tid; e < kH * kIn; e += kTrainThreads) {
      const int j = e / kIn, f = e % kIn;
      float s = 0.0f;
      for (int i = 0; i < B; ++i) s = fmaf(H1[i * kH + j], Xb[i * kIn + f], s);
      G[kOffW1 + e] = s;
    }
    if (tid < kH) {
      float s = 0.0f;
      for (int i = 0; i < B; ++i) s += H1[i * kH + tid];
      G[kOffB1 + tid] = s;
    }
    __syncthreads();
    if (a.grad_out) {
      for (int i = tid; i < kNParams; i += kTrainThreads) a.grad_out[i] = G[i];
      return;
    }
    // Adam
    const int t = a.t0 + step + 1;
    const float c1 = 1.0f - powf(a.beta1, (float)t);
    const float c2 = 1.0f - powf(a.beta2, (float)t);
    for (int i = tid; i < kNParams; i += kTrainThreads) {
      const float gi = G[i];
      const float m = a.beta1 * a.adam_m[i] + (1.0f - a.beta1) * gi;
      const float v = a.beta2 * a.adam_v[i] + (1.0f - a.beta2) * gi * gi;
      a.adam_m[i] = m;
      a.adam_v[i] = v;
      P[i] -= a.lr * (m / c1) / (sqrtf(v / c2) + a.eps);
    }
    __syncthreads();
  }
  for (int i = tid; i < kNParams; i += kTrainThreads) a.params[i] = P[i];
}

static size_t train_smem() {
  return sizeof(float) * (2 * kNParams + kMaxBatch * kIn + 2 * kMaxBatch * kH + kMaxBatch + 256);
}

}  // namespace tla

using namespace tla;

extern "C" int tla_mlp_num_params(void) { return kNParams; }

extern "C" int tla_mlp_forward(const float* params, const float* X, int n, float* out, void* stream) {
  TLA_REQUIRE(params && X && out, "mlp_forward: NULL pointer");
  TLA_REQUIRE(n >= 0, "mlp_forward: n = %d", n);
  if (n == 0) return TLA_OK;
  mlp_forward_kernel<<<ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(params, X, n, out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" int tla_mlp_train(float* params, float* adam_m, float* adam_v, const float* X, const float* y,
                             const int* order, int n, int batch, int steps, float lr, float beta1, float beta2,
                             float eps, int t0, float* loss, float* grad_out, void* stream) {
  TLA_REQUIRE(params && X && y && order, "mlp_train: NULL pointer");
  TLA_REQUIRE(grad_out || (adam_m && adam_v), "mlp_train: Adam state is NULL");
  TLA_REQUIRE(n >= 1 && batch >= 1 && batch <= kMaxBatch, "mlp_train: n = %d, batch = %d (max %d)", n, batch,
              kMaxBatch);
  TLA_REQUIRE(steps >= 0, "mlp_train: steps = %d", steps);
  if (steps == 0) return TLA_OK;
  TrainArgs a{params, adam_m, adam_v, X, y, order, n, batch, steps, lr, beta1, beta2, eps, t0, loss, grad_out};
  const size_t smem = train_smem();
  TLA_HIP(hipFuncSetAttribute((const void*)mlp_train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem));
  mlp_train_kernel<<<1, kTrainThreads, smem, (hipStream_t)stream>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
