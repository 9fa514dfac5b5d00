// Reduced QR of a tall matrix by Householder reflections, the factorisation
// `tl.qr` computes for parafac(..., orthogonalise=...) (tensorly/
// decomposition/_cp.py:395-399; numpy's qr: LAPACK geqrf + orgqr).
//
// One workgroup per matrix, FP64 arithmetic on a column-major copy in the
// workspace.  Reflector k is LAPACK's: for alpha = A[k,k] and x = A[k+1:, k],
// beta = -sign(alpha) sqrt(alpha^2 + ||x||^2), tau = (beta - alpha) / beta,
// v = x / (alpha - beta) (tau = 0 and no reflection when x = 0), so the
// signs of Q's columns and R's diagonal are LAPACK's.  Q is then formed by
// accumulating the reflectors backwards as orgqr does.
#include <math.h>

#include "common.h"

namespace tla {

constexpr int kQrThreads = 512;

struct QrArgs {
  const void* A;  // rows x cols, row-major (dtype)
  void* Q;        // rows x cols, row-major (dtype), may alias A
  void* Rm;       // cols x cols upper triangular (dtype), nullable
  double* W;      // rows x cols FP64 workspace, column-major
  double* tau;    // cols
  int64_t rows;
  int cols;
};

__device__ double qr_block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  return s;
}

template <typename T>
__global__ __launch_bounds__(kQrThreads) void householder_qr_kernel(QrArgs a) {
  __shared__ double red[kQrThreads / 64];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
  const int64_t m = a.rows;
  const int n = a.cols;
  double* W = a.W;
  for (int64_t e = tid; e < m * n; e += nt) {
    const int64_t i = e / n;
    const int j = (int)(e % n);
    W[(int64_t)j * m + i] = (double)((const T*)a.A)[e];
  }
  __syncthreads();
  // factorisation: reflector k, then its action on the columns right of k
  for (int k = 0; k < n; ++k) {
    double* wk = W + (int64_t)k * m;
    double s = 0.0;
    for (int64_t i = k + 1 + tid; i < m; i += nt) s = fma(wk[i], wk[i], s);
    s = qr_block_sum(s, red);
    const double alpha = wk[k];
    double tk = 0.0;
    if (s > 0.0) {
      const double beta = -copysign(sqrt(fma(alpha, alpha, s)), alpha);
      tk = (beta - alpha) / beta;
      const double scal = 1.0 / (alpha - beta);
      for (int64_t i = k + 1 + tid; i < m; i += nt) wk[i] *= scal;
      __syncthreads();
      if (tid == 0) wk[k] = beta;
    }
    if (tid == 0) a.tau[k] = tk;
    __syncthreads();
    if (tk != 0.0) {
      for (int j = k + 1 + wave; j < n; j += nw) {
        double* wj = W + (int64_t)j * m;
        double d = 0.0;
        for (int64_t i = k + 1 + lane; i < m; i += 64) d = fma(wk[i], wj[i], d);
        d = wave_sum(d) + wj[k];
        const double f = tk * d;
        if (lane == 0) wj[k] -= f;
        for (int64_t i = k + 1 + lane; i < m; i += 64) wj[i] = fma(-f, wk[i], wj[i]);
      }
    }
    __syncthreads();
  }
  if (a.Rm) {
    for (int e = tid; e < n * n; e += nt) {
      const int r = e / n, c = e % n;
      ((T*)a.Rm)[e] = (T)(c >= r ? W[(int64_t)c * m + r] : 0.0);
    }
  }
  __syncthreads();
  // Q = H_0 ... H_{n-1} [I; 0], from the last reflector back (orgqr): when
  // reflector i is applied, columns j > i are already columns of Q and are
  // zero in rows < j
  for (int i = n - 1; i >= 0; --i) {
    double* wi = W + (int64_t)i * m;
    const double ti = a.tau[i];
    for (int j = i + 1 + wave; j < n; j += nw) {
      double* wj = W + (int64_t)j * m;
      double d = 0.0;
      for (int64_t r = i + 1 + lane; r < m; r += 64) d = fma(wi[r], wj[r], d);
      d = wave_sum(d) + wj[i];
      const double f = ti * d;
      if (lane == 0) wj[i] -= f;
      for (int64_t r = i + 1 + lane; r < m; r += 64) wj[r] = fma(-f, wi[r], wj[r]);
    }
    __syncthreads();
    for (int64_t r = i + 1 + tid; r < m; r += nt) wi[r] *= -ti;
    for (int64_t r = tid; r < i; r += nt) wi[r] = 0.0;
    if (tid == 0) wi[i] = 1.0 - ti;
    __syncthreads();
  }
  for (int64_t e = tid; e < m * n; e += nt) {
    const int64_t i = e / n;
    const int j = (int)(e % n);
    ((T*)a.Q)[e] = (T)W[(int64_t)j * m + i];
  }
}

}  // namespace tla

using namespace tla;

extern "C" size_t tla_qr_workspace_size(int64_t rows, int cols) {
  if (rows < 1 || cols < 1) return 0;
  return (size_t(rows) * cols + size_t(cols)) * sizeof(double);
}

extern "C" int tla_qr(int dtype, const void* A, int64_t rows, int cols, void* Q, void* Rmat, void* workspace,
                      size_t workspace_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "qr: bad dtype %d", dtype);
  TLA_REQUIRE(cols >= 1 && rows >= cols, "qr: shape (%lld, %d) must have rows >= cols >= 1", (long long)rows, cols);
  TLA_REQUIRE(A && Q && workspace, "qr: NULL pointer");
  TLA_REQUIRE(workspace_bytes >= tla_qr_workspace_size(rows, cols), "qr: workspace too small");
  QrArgs a{A, Q, Rmat, (double*)workspace, (double*)workspace + size_t(rows) * cols, rows, cols};
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) householder_qr_kernel<float><<<1, kQrThreads, 0, st>>>(a);
  else householder_qr_kernel<double><<<1, kQrThreads, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
