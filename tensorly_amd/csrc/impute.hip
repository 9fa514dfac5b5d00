// Masked imputation for parafac(..., mask=...) (tensorly/decomposition/
// _cp.py:194-207 and :437-445): the missing entries of the working tensor
// are replaced by the current CP model,
//   X <- X * mask + [[w; A_0, ..., A_{N-1}]] * (1 - mask),
// in one pass that also sums ||X_new||^2 and ||X_new - model||^2 (the
// reference's norm_tensor and reconstruction error).  The model is never
// stored: each workgroup takes 64 (float) or 32 (double) rows of the C-order
// tensor (all modes but the last), forms their rank-wise scales s_r = w_r prod_{m < N-1}
// A_m[i_m, r] in LDS, and streams the last mode in chunks of 64 whose factor
// rows are staged transposed in LDS, so a lane's model entry is R
// multiply-adds of LDS operands.  Partial sums per workgroup, then a fixed
// order final sum: repeatable bit for bit.
//
// tla_masked_blend is the same blend for a dense matrix and a given
// low-rank approximation (the mask imputation of tensorly/tenalg/svd.py:
// 431-439).
#include "common.h"

namespace tla {

// rows per wave: 16 (float) or 8 (double, keeping the LDS under 64 KB at
// rank 64); four waves per workgroup
template <typename T>
__host__ __device__ constexpr int imp_rows_per_wave() { return sizeof(T) == 8 ? 8 : 16; }

struct ImputeArgs {
  void* X;           // P x K (dtype), in / out
  const void* mask;  // P x K (dtype)
  const void* fac[kMaxOrder];
  int64_t sz[kMaxOrder];
  int nlead;  // modes forming the row index (ndim - 1)
  const void* weights;  // R (dtype) or NULL
  int R;
  int64_t P, K;
  int write;
  double* partial;  // [gridDim.x][2]
};

template <typename T, int RT>
__global__ __launch_bounds__(256) void cp_impute_kernel(ImputeArgs a) {
  constexpr int RPW = imp_rows_per_wave<T>();
  __shared__ T Ct[RT][65];
  __shared__ T S[4][RPW][RT];
  __shared__ double red[2][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int R = a.R;
  const int64_t P = a.P, K = a.K;
  const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * RPW;
  const T* C = (const T*)a.fac[a.nlead];
  for (int j = 0; j < RPW; ++j) {
    const int64_t q = q0 + j;
    T s = T(0);
    if (q < P && lane < R) {
      s = a.weights ? ((const T*)a.weights)[lane] : T(1);
      int64_t rem = q;
      for (int m = a.nlead - 1; m >= 0; --m) {
        const int64_t im = rem % a.sz[m];
        rem /= a.sz[m];
        s *= ((const T*)a.fac[m])[im * R + lane];
      }
    }
    if (lane < RT) S[wave][j][lane] = s;
  }
  double sx = 0.0, sd = 0.0;
  T* X = (T*)a.X;
  const T* Mk = (const T*)a.mask;
  for (int64_t k0 = 0; k0 < K; k0 += 64) {
    __syncthreads();
    for (int e = tid; e < 64 * RT; e += 256) {
      const int kk = e / RT, r = e % RT;
      const int64_t k = k0 + kk;
      Ct[r][kk] = (k < K && r < R) ? C[k * R + r] : T(0);
    }
    __syncthreads();
    const int64_t k = k0 + lane;
    if (k < K) {
      for (int j = 0; j < RPW; ++j) {
        const int64_t q = q0 + j;
        if (q >= P) break;
        T rec = T(0);
#pragma unroll
        for (int r = 0; r < RT; ++r) rec = fma(S[wave][j][r], Ct[r][lane], rec);
        const int64_t idx = q * K + k;
        const T m = Mk[idx];
        const T xn = X[idx] * m + rec * (T(1) - m);
        const double d = (double)xn - (double)rec;
        sx = fma((double)xn, (double)xn, sx);
        sd = fma(d, d, sd);
        if (a.write) X[idx] = xn;
      }
    }
  }
  sx = wave_sum(sx);
  sd = wave_sum(sd);
  if (lane == 0) {
    red[0][wave] = sx;
    red[1][wave] = sd;
  }
  __syncthreads();
  if (tid == 0) {
    a.partial[2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    a.partial[2 * blockIdx.x + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// sums[c] = sum over blocks b of partial[2 b + c], in block order per lane
// then lanes combined in a fixed tree.
__global__ __launch_bounds__(64) void impute_sums_kernel(const double* partial, int nblk, double* sums) {
  const int lane = threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  for (int b = lane; b < nblk; b += 64) {
    s0 += partial[2 * b];
    s1 += partial[2 * b + 1];
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (lane == 0) {
    sums[0] = s0;
    sums[1] = s1;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void masked_blend_kernel(const T* X, const T* mask, const T* L, T* out, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const T m = mask[e];
    out[e] = X[e] * m + L[e] * (T(1) - m);
  }
}

}  // namespace tla

using namespace tla;

// workgroups of a launch: four waves of imp_rows_per_wave rows each
static int64_t impute_blocks(int ndim, const int64_t* shape, int dtype) {
  int64_t P = 1;
  for (int m = 0; m < ndim - 1; ++m) P *= shape[m];
  const int64_t rows = 4 * (dtype == TLA_F64 ? imp_rows_per_wave<double>() : imp_rows_per_wave<float>());
  return (P + rows - 1) / rows;
}

extern "C" size_t tla_cp_impute_workspace_size(int dtype, int ndim, const int64_t* shape) {
  if (ndim < 2 || ndim > kMaxOrder || !shape) return 0;
  return size_t(impute_blocks(ndim, shape, dtype)) * 2 * sizeof(double);
}

extern "C" int tla_cp_impute(int dtype, void* X, const void* mask, int ndim, const int64_t* shape,
                             const void* const* factors, const void* weights, int rank, int write, double* sums,
                             void* workspace, size_t workspace_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "cp_impute: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "cp_impute: order %d not in [2, %d]", ndim, kMaxOrder);
  TLA_REQUIRE(rank >= 1 && rank <= kMaxRank, "cp_impute: rank %d not in [1, %d]", rank, kMaxRank);
  TLA_REQUIRE(X && mask && factors && sums && workspace, "cp_impute: NULL pointer");
  TLA_REQUIRE(workspace_bytes >= tla_cp_impute_workspace_size(dtype, ndim, shape), "cp_impute: workspace too small");
  ImputeArgs a{};
  for (int m = 0; m < ndim; ++m) {
    TLA_REQUIRE(shape[m] >= 1 && factors[m], "cp_impute: mode %d empty or without factor", m);
    a.fac[m] = factors[m];
    a.sz[m] = shape[m];
  }
  a.X = X; a.mask = mask; a.nlead = ndim - 1; a.weights = weights; a.R = rank;
  a.P = 1;
  for (int m = 0; m < ndim - 1; ++m) a.P *= shape[m];
  a.K = shape[ndim - 1];
  a.write = write;
  a.partial = (double*)workspace;
  const int64_t nblk = impute_blocks(ndim, shape, dtype);
  TLA_REQUIRE(nblk <= 0x7fffffff, "cp_impute: too many rows");
  hipStream_t st = (hipStream_t)stream;
  const int RT = rank <= 16 ? 16 : (rank <= 32 ? 32 : (rank <= 48 ? 48 : 64));
#define TLA_IMP_CASE(RTV)                                                                          \
  case RTV:                                                                                        \
    if (dtype == TLA_F32) cp_impute_kernel<float, RTV><<<(unsigned)nblk, 256, 0, st>>>(a);         \
    else cp_impute_kernel<double, RTV><<<(unsigned)nblk, 256, 0, st>>>(a);                         \
    break;
  switch (RT) {
    TLA_IMP_CASE(16)
    TLA_IMP_CASE(32)
    TLA_IMP_CASE(48)
    TLA_IMP_CASE(64)
  }
#undef TLA_IMP_CASE
  TLA_LAUNCH_CHECK();
  impute_sums_kernel<<<1, 64, 0, st>>>(a.partial, (int)nblk, sums);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" int tla_masked_blend(int dtype, const void* X, const void* mask, const void* low_rank, void* out, int64_t n,
                                void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "masked_blend: bad dtype %d", dtype);
  TLA_REQUIRE(n >= 0, "masked_blend: n < 0");
  if (n == 0) return TLA_OK;
  TLA_REQUIRE(X && mask && low_rank && out, "masked_blend: NULL pointer");
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32)
    masked_blend_kernel<float><<<(unsigned)blocks, 256, 0, st>>>((const float*)X, (const float*)mask,
                                                                 (const float*)low_rank, (float*)out, n);
  else
    masked_blend_kernel<double><<<(unsigned)blocks, 256, 0, st>>>((const double*)X, (const double*)mask,
                                                                  (const double*)low_rank, (double*)out, n);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
