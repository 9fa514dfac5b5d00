// FP64 pieces of the ALS sweep: factor Grams, the Hadamard-of-Grams normal
// equations with a Cholesky solve (LU fallback), the squared norm of the
// tensor and the reconstruction error.  Everything here is R x R or I x R
// sized -- launch/latency bound -- so each step is one small fused kernel and
// the R x R work is done redundantly per workgroup rather than paying a
// second launch.
#include <math.h>

#include "common.h"

namespace tla {

constexpr int kMaxRank = 64;
constexpr int kSolveRows = 64;  // factor rows handled by one solve workgroup

template <typename T>
__device__ __forceinline__ double ld(const void* p, int64_t i) {
  return (double)((const T*)p)[i];
}

// ---------------------------------------------------------------------------
// Squared norm: per-block FP64 partials, then one block sums them in order.
// ---------------------------------------------------------------------------
constexpr int kNormBlocks = 1024;

template <typename T>
__global__ __launch_bounds__(256) void sqnorm_partial_kernel(const T* __restrict__ x, int64_t n,
                                                             double* __restrict__ partial) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = (double)x[i];
    s = fma(v, v, s);
  }
  s = wave_sum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  s = wave_sum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// Gram G = A^T A in FP64: blocks of kSolveRows rows give partial R x R
// matrices, summed in block order by gram_finalize_kernel.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void gram_partial_kernel(const T* __restrict__ A, int64_t rows, int R,
                                                           double* __restrict__ partial) {
  __shared__ double tile[kSolveRows * kMaxRank];
  const int64_t r0 = (int64_t)blockIdx.x * kSolveRows;
  const int nr = (int)min((int64_t)kSolveRows, rows - r0);
  for (int e = threadIdx.x; e < nr * R; e += 256) tile[e] = (double)A[r0 * R + e];
  __syncthreads();
  for (int e = threadIdx.x; e < R * R; e += 256) {
    const int r = e / R, s = e % R;
    double acc = 0.0;
    for (int i = 0; i < nr; ++i) acc = fma(tile[i * R + r], tile[i * R + s], acc);
    partial[(int64_t)blockIdx.x * R * R + e] = acc;
  }
}

__global__ __launch_bounds__(256) void gram_finalize_kernel(const double* __restrict__ partial, int nblk,
                                                            int R, double* __restrict__ G,
                                                            const double* __restrict__ ip_partial,
                                                            double* __restrict__ iprod) {
  for (int e = threadIdx.x; e < R * R; e += 256) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * R * R + e];
    G[e] = s;
  }
  if (iprod != nullptr && threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += ip_partial[b];
    *iprod = s;
  }
}

// ---------------------------------------------------------------------------
// Normal-equations solve.  Every block rebuilds V, factors it and forms
// V^-1 in LDS (R <= 64: a few microseconds), then handles kSolveRows rows of
// the new factor, its partial Gram and partial <M, A>.
// ---------------------------------------------------------------------------
struct SolveArgs {
  const double* grams[kMaxOrder];
  int ndim, mode, R;
  const void* weights;
  double l2;
  const void* mttkrp;
  int64_t rows;
  void* factor;
  double* gram_partial;  // [nblk][R][R]
  double* ip_partial;    // [nblk] or nullptr
  int* status;           // nullable
  int* counter;          // nullable: zero-initialised block counter -> the last block finalises
  double* gram_out;
  double* iprod_out;
};

// In-place Cholesky of the R x R matrix in LDS (lower triangle).  Returns
// false if a pivot is not positive and finite.
__device__ bool lds_cholesky(double* V, int R, int* flag) {
  const int tid = threadIdx.x;
  if (tid == 0) *flag = 0;
  __syncthreads();
  for (int k = 0; k < R; ++k) {
    if (tid == 0) {
      const double d = V[k * R + k];
      if (!(d > 0.0) || !isfinite(d)) *flag = 1;
      V[k * R + k] = sqrt(fmax(d, 0.0));
    }
    __syncthreads();
    if (*flag) return false;
    const double dk = V[k * R + k];
    for (int i = k + 1 + tid; i < R; i += blockDim.x) V[i * R + k] /= dk;
    __syncthreads();
    const int m = R - k - 1;
    for (int e = tid; e < m * m; e += blockDim.x) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      if (j <= i) V[i * R + j] -= V[i * R + k] * V[j * R + k];
    }
    __syncthreads();
  }
  return true;
}

// Inverse of L L^T into Vi (full matrix): thread s solves L L^T x = e_s.
__device__ void lds_chol_inverse(const double* L, double* Vi, int R) {
  const int s = threadIdx.x;
  if (s < R) {
    // forward: L y = e_s, stored in column s of Vi
    for (int i = 0; i < R; ++i) {
      double v = (i == s) ? 1.0 : 0.0;
      for (int j = 0; j < i; ++j) v -= L[i * R + j] * Vi[j * R + s];
      Vi[i * R + s] = v / L[i * R + i];
    }
    // backward: L^T x = y
    for (int i = R - 1; i >= 0; --i) {
      double v = Vi[i * R + s];
      for (int j = i + 1; j < R; ++j) v -= L[j * R + i] * Vi[j * R + s];
      Vi[i * R + s] = v / L[i * R + i];
    }
  }
  __syncthreads();
}

// Gauss-Jordan inverse with partial pivoting (fallback for V that is not
// numerically positive definite).  A is destroyed.  Returns false if singular.
__device__ bool lds_gj_inverse(double* A, double* Vi, int R, int* piv_shared, int* flag) {
  const int tid = threadIdx.x;
  for (int e = tid; e < R * R; e += blockDim.x) Vi[e] = ((e / R) == (e % R)) ? 1.0 : 0.0;
  if (tid == 0) *flag = 0;
  __syncthreads();
  for (int k = 0; k < R; ++k) {
    if (tid == 0) {
      int p = k;
      double best = fabs(A[k * R + k]);
      for (int i = k + 1; i < R; ++i) {
        const double v = fabs(A[i * R + k]);
        if (v > best) { best = v; p = i; }
      }
      *piv_shared = p;
      if (!(best > 0.0) || !isfinite(best)) *flag = 1;
    }
    __syncthreads();
    if (*flag) return false;
    const int p = *piv_shared;
    if (p != k) {
      for (int j = tid; j < R; j += blockDim.x) {
        double t = A[k * R + j]; A[k * R + j] = A[p * R + j]; A[p * R + j] = t;
        t = Vi[k * R + j]; Vi[k * R + j] = Vi[p * R + j]; Vi[p * R + j] = t;
      }
    }
    __syncthreads();
    const double inv = 1.0 / A[k * R + k];
    __syncthreads();
    for (int j = tid; j < R; j += blockDim.x) { A[k * R + j] *= inv; Vi[k * R + j] *= inv; }
    __syncthreads();
    for (int e = tid; e < R * R; e += blockDim.x) {
      const int i = e / R, j = e % R;
      if (i != k) {
        const double f = A[i * R + k];
        if (j != k) A[i * R + j] -= f * A[k * R + j];
        Vi[i * R + j] -= f * Vi[k * R + j];
      }
    }
    __syncthreads();
    for (int i = tid; i < R; i += blockDim.x)
      if (i != k) A[i * R + k] = 0.0;
    __syncthreads();
  }
  return true;
}

template <typename T>
__device__ void build_v(const SolveArgs& a, double* V) {
  const int R = a.R;
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    const int r = e / R, s = e % R;
    double v = 1.0;
    for (int m = 0; m < a.ndim; ++m)
      if (m != a.mode) v *= a.grams[m][e];
    if (r == s) v += a.l2;
    if (a.weights) v *= ld<T>(a.weights, r) * ld<T>(a.weights, s);
    V[e] = v;
  }
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(256) void cp_solve_kernel(SolveArgs a) {
  __shared__ double Lbuf[kMaxRank * kMaxRank];
  __shared__ double Vi[kMaxRank * kMaxRank];
  __shared__ double Mr[kSolveRows * kMaxRank];
  __shared__ int flag, piv, ticket;
  const int R = a.R, tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kSolveRows;
  const int nr = (int)min((int64_t)kSolveRows, a.rows - r0);
  // this block's MTTKRP rows, staged once (coalesced)
  for (int e = tid; e < nr * R; e += blockDim.x) Mr[e] = ld<T>(a.mttkrp, r0 * R + e);
  build_v<T>(a, Lbuf);
  int status = 0;
  if (!lds_cholesky(Lbuf, R, &flag)) {
    build_v<T>(a, Lbuf);
    status = lds_gj_inverse(Lbuf, Vi, R, &piv, &flag) ? 1 : 2;
  } else {
    lds_chol_inverse(Lbuf, Vi, R);
  }
  if (blockIdx.x == 0 && tid == 0 && a.status) *a.status = status;
  __syncthreads();
  // New rows: A[i, s] = sum_r M[i, r] Vi[r, s]  (V symmetric); keep the
  // rounded values in LDS (reusing Lbuf) for the Gram / inner product.
  double* An = Lbuf;  // kSolveRows x R
  double ip = 0.0;
  for (int e = tid; e < nr * R; e += blockDim.x) {
    const int i = e / R, s = e % R;
    const double* mrow = Mr + i * R;
    double v0 = 0.0, v1 = 0.0;
    int r = 0;
    for (; r + 1 < R; r += 2) {
      v0 = fma(mrow[r], Vi[r * R + s], v0);
      v1 = fma(mrow[r + 1], Vi[(r + 1) * R + s], v1);
    }
    if (r < R) v0 = fma(mrow[r], Vi[r * R + s], v0);
    const T out = (T)(v0 + v1);
    ((T*)a.factor)[(r0 + i) * R + s] = out;
    An[i * R + s] = (double)out;
    ip = fma(mrow[s], (double)out, ip);
  }
  __syncthreads();
  for (int e = tid; e < R * R; e += blockDim.x) {
    const int r = e / R, s = e % R;
    double acc0 = 0.0, acc1 = 0.0;
    int i = 0;
    for (; i + 1 < nr; i += 2) {
      acc0 = fma(An[i * R + r], An[i * R + s], acc0);
      acc1 = fma(An[(i + 1) * R + r], An[(i + 1) * R + s], acc1);
    }
    if (i < nr) acc0 = fma(An[i * R + r], An[i * R + s], acc0);
    a.gram_partial[(int64_t)blockIdx.x * R * R + e] = acc0 + acc1;
  }
  if (a.ip_partial) {
    ip = wave_sum(ip);
    __shared__ double red[4];
    if ((tid & 63) == 0) red[tid >> 6] = ip;
    __syncthreads();
    if (tid == 0) a.ip_partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
  }
  if (!a.counter) return;
  // The last block to finish sums every block's partials in block order.
  __threadfence();
  __syncthreads();
  if (tid == 0) ticket = atomicAdd(a.counter, 1);
  __syncthreads();
  if (ticket != (int)gridDim.x - 1) return;
  __threadfence();
  const int nblk = gridDim.x;
  for (int e = tid; e < R * R; e += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b)
      s += __hip_atomic_load(a.gram_partial + (int64_t)b * R * R + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a.gram_out[e] = s;
  }
  if (a.ip_partial && tid == 0) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += __hip_atomic_load(a.ip_partial + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *a.iprod_out = s;
  }
  if (tid == 0) *a.counter = 0;
}

// err = sqrt(|normx2 + sum((prod G) * w w^T) - 2 iprod|); errors[slot] = err / sqrt(normx2)
template <typename T>
__global__ __launch_bounds__(256) void cp_error_kernel(SolveArgs a, const double* normx2, const double* iprod,
                                                       double* errors, int slot) {
  const int R = a.R;
  double s = 0.0;
  for (int e = threadIdx.x; e < R * R; e += 256) {
    const int r = e / R, q = e % R;
    double v = 1.0;
    for (int m = 0; m < a.ndim; ++m) v *= a.grams[m][e];
    if (a.weights) v *= ld<T>(a.weights, r) * ld<T>(a.weights, q);
    s += v;
  }
  s = wave_sum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mn2 = (red[0] + red[1]) + (red[2] + red[3]);
    const double nx2 = *normx2;
    const double err = sqrt(fabs(nx2 + mn2 - 2.0 * (*iprod)));
    errors[slot] = err / sqrt(nx2);
  }
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace tla

using namespace tla;

extern "C" size_t tla_sqnorm_workspace_size(int64_t n) {
  (void)n;
  return align256(kNormBlocks * sizeof(double));
}

extern "C" int tla_sqnorm(int dtype, const void* X, int64_t n, double* out, void* ws, size_t ws_bytes,
                          void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "sqnorm: bad dtype %d", dtype);
  TLA_REQUIRE(ws_bytes >= tla_sqnorm_workspace_size(n), "sqnorm: workspace too small");
  TLA_REQUIRE(out && (X || n == 0), "sqnorm: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  double* partial = (double*)ws;
  const int nblk = (int)std::min<int64_t>(kNormBlocks, std::max<int64_t>(1, (n + 255) / 256));
  if (dtype == TLA_F32)
    sqnorm_partial_kernel<float><<<nblk, 256, 0, st>>>((const float*)X, n, partial);
  else
    sqnorm_partial_kernel<double><<<nblk, 256, 0, st>>>((const double*)X, n, partial);
  TLA_LAUNCH_CHECK();
  sum_partials_kernel<<<1, 256, 0, st>>>(partial, nblk, out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" size_t tla_gram64_workspace_size(int64_t rows, int rank) {
  const int64_t nblk = (rows + kSolveRows - 1) / kSolveRows;
  return align256(size_t(nblk) * rank * rank * sizeof(double)) + align256(size_t(nblk) * sizeof(double));
}

extern "C" int tla_gram64(int dtype, const void* A, int64_t rows, int rank, double* G, void* ws,
                          size_t ws_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "gram64: bad dtype %d", dtype);
  TLA_REQUIRE(rank >= 1 && rank <= kMaxRank, "gram64: rank %d not in [1, %d]", rank, kMaxRank);
  TLA_REQUIRE(rows >= 1, "gram64: empty factor");
  TLA_REQUIRE(ws_bytes >= tla_gram64_workspace_size(rows, rank), "gram64: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  const int nblk = (int)((rows + kSolveRows - 1) / kSolveRows);
  double* partial = (double*)ws;
  if (dtype == TLA_F32)
    gram_partial_kernel<float><<<nblk, 256, 0, st>>>((const float*)A, rows, rank, partial);
  else
    gram_partial_kernel<double><<<nblk, 256, 0, st>>>((const double*)A, rows, rank, partial);
  TLA_LAUNCH_CHECK();
  gram_finalize_kernel<<<1, 256, 0, st>>>(partial, nblk, rank, G, nullptr, nullptr);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" size_t tla_cp_solve_workspace_size(int64_t rows, int rank) {
  return tla_gram64_workspace_size(rows, rank);
}

extern "C" int tla_cp_solve(int dtype, const double* const* grams, int ndim, int mode,
                            const void* weights, double l2_reg, const void* mttkrp, int64_t rows,
                            int rank, void* factor_out, double* gram_out, double* iprod_out,
                            int* status_out, int* counter, void* ws, size_t ws_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "cp_solve: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "cp_solve: order %d", ndim);
  TLA_REQUIRE(mode >= 0 && mode < ndim, "cp_solve: mode %d", mode);
  TLA_REQUIRE(rank >= 1 && rank <= kMaxRank, "cp_solve: rank %d not in [1, %d]", rank, kMaxRank);
  TLA_REQUIRE(rows >= 1, "cp_solve: empty factor");
  TLA_REQUIRE(ws_bytes >= tla_cp_solve_workspace_size(rows, rank), "cp_solve: workspace too small");
  TLA_REQUIRE(mttkrp && factor_out && gram_out, "cp_solve: NULL pointer");
  SolveArgs a{};
  for (int m = 0; m < ndim; ++m) {
    a.grams[m] = grams[m];
    TLA_REQUIRE(m == mode || grams[m] != nullptr, "cp_solve: gram %d is NULL", m);
  }
  a.ndim = ndim; a.mode = mode; a.R = rank;
  a.weights = weights; a.l2 = l2_reg; a.mttkrp = mttkrp; a.rows = rows; a.factor = factor_out;
  const int nblk = (int)((rows + kSolveRows - 1) / kSolveRows);
  a.gram_partial = (double*)ws;
  a.ip_partial = iprod_out ? (double*)((char*)ws + align256(size_t(nblk) * rank * rank * 8)) : nullptr;
  a.status = status_out;
  a.counter = counter;
  a.gram_out = gram_out;
  a.iprod_out = iprod_out;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) cp_solve_kernel<float><<<nblk, 256, 0, st>>>(a);
  else cp_solve_kernel<double><<<nblk, 256, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  if (!counter) {
    gram_finalize_kernel<<<1, 256, 0, st>>>(a.gram_partial, nblk, rank, gram_out, a.ip_partial, iprod_out);
    TLA_LAUNCH_CHECK();
  }
  return TLA_OK;
}

extern "C" int tla_cp_error(const double* const* grams, int ndim, const void* weights, int dtype,
                            int rank, const double* normx2, const double* iprod, double* errors,
                            int slot, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "cp_error: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 1 && ndim <= kMaxOrder, "cp_error: order %d", ndim);
  TLA_REQUIRE(normx2 && iprod && errors, "cp_error: NULL pointer");
  SolveArgs a{};
  for (int m = 0; m < ndim; ++m) a.grams[m] = grams[m];
  a.ndim = ndim; a.R = rank; a.weights = weights; a.mode = -1;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) cp_error_kernel<float><<<1, 256, 0, st>>>(a, normx2, iprod, errors, slot);
  else cp_error_kernel<double><<<1, 256, 0, st>>>(a, normx2, iprod, errors, slot);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
