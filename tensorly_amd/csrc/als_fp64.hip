// FP64 pieces of the ALS sweep: factor Grams, the Hadamard-of-Grams normal
// equations solved by symmetric positive definite elimination (LU with
// partial pivoting as the fallback), the squared norm of the
// tensor and the reconstruction error.  Everything here is R x R or I x R
// sized -- launch/latency bound -- so each step is one small fused kernel and
// the R x R work is done redundantly per workgroup rather than paying a
// second launch.
#include <math.h>

#include "common.h"

namespace tla {

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
// Normal-equations solve.  Every block rebuilds V and inverts it in
// registers (R <= 64: a few microseconds), then handles kSolveRows rows of
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
  // optional, last mode only: errors[err_slot] = relative error of the model
  // after this update (needs iprod_out), computed by the finalising block
  const double* normx2;
  double* errors;
  int err_slot;
  // optional: V^-1 computed ahead (R x R FP64, global) and its status
  const double* vi_pre;
  const int* vi_status;
};

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}


// V^-1 of the symmetric positive definite V (R x R in LDS, left intact) by
// Gauss-Jordan without pivoting,
// held in registers: wave w owns rows w, w + 4, ... of [V | I], lane l owns
// columns l and l + 64.  Per elimination step the owner of the pivot row
// publishes it through LDS (double-buffered, one barrier per step) and every
// row reads its multiplier from the pivot column with a lane read.  For SPD V
// the pivots are the (positive) Schur complement diagonals; a pivot that is
// not positive and finite returns false (the caller falls back to partial
// pivoting).  Needs blockDim.x == 256 and R <= 64.
__device__ __forceinline__ bool spd_gj_inverse(const double* V, int R, double* Vi, double* pbuf, int* flag) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nq = (R - w + 3) >> 2;
  double x[16][2];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = w + 4 * q;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = lane + 64 * h;
      double v = 0.0;
      if (q < nq) v = (c < R) ? V[i * R + c] : ((c - R == i) ? 1.0 : 0.0);
      x[q][h] = v;
    }
  }
  for (int k = 0; k < R; ++k) {
    const int qk = k >> 2, kl = k & 63, kh = k >> 6;
    double* buf = pbuf + (k & 1) * 128;
    if (w == (k & 3)) {
      double r0 = 0.0, r1 = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q == qk) {
          r0 = x[q][0];
          r1 = x[q][1];
        }
      const double piv = readlane_f64(kh ? r1 : r0, kl);
      const bool bad = !(piv > 0.0) || !isfinite(piv);
      const double inv = 1.0 / piv;
      r0 *= inv;
      r1 *= inv;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q == qk) {
          x[q][0] = r0;
          x[q][1] = r1;
        }
      buf[lane] = r0;
      buf[64 + lane] = r1;
      if (lane == 0) flag[k & 1] = bad ? 1 : 0;
    }
    __syncthreads();
    if (flag[k & 1]) return false;
    const double p0 = buf[lane], p1 = buf[64 + lane];
    // branch-free over all 16 row slots: rows outside [0, R) hold zeros, the
    // pivot row gets a zero multiplier (uniform selects, no divergent jumps)
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const double fr = readlane_f64(kh ? x[q][1] : x[q][0], kl);
      const double f = (w + 4 * q != k) ? fr : 0.0;
      x[q][0] = fma(-f, p0, x[q][0]);
      x[q][1] = fma(-f, p1, x[q][1]);
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = w + 4 * q;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = lane + 64 * h;
      if (q < nq && c >= R && c < 2 * R) Vi[i * R + (c - R)] = x[q][h];
    }
  }
  __syncthreads();
  return true;
}

// Same elimination for R <= 32 inside ONE wave (lane c owns column c of
// [V | I], rows in registers): no barriers and no LDS traffic per pivot, which
// makes the R sequential steps cheap.  Called by wave 0 only; Vi and the
// return value are valid after the caller's next barrier / in wave 0.
__device__ __forceinline__ bool wave_spd_inverse32(const double* V, int R, double* Vi) {
  const int lane = threadIdx.x & 63;
  double x[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    double v = 0.0;
    if (i < R) v = (lane < R) ? V[i * R + lane] : ((lane - R == i) ? 1.0 : 0.0);
    x[i] = v;
  }
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    if (k < R) {
      const double piv = readlane_f64(x[k], k);
      ok = ok && (piv > 0.0) && isfinite(piv);
      x[k] *= 1.0 / piv;
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        if (i != k) {
          const double f = readlane_f64(x[i], k);
          x[i] = fma(-f, x[k], x[i]);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 32; ++i)
    if (i < R && lane >= R && lane < 2 * R) Vi[i * R + (lane - R)] = x[i];
  return ok;
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

// V (built from the Grams into Vbuf) and V^-1 into Vi, both LDS R x R.
// Returns 0 (SPD elimination), 1 (partial-pivoting fallback) or 2 (singular).
template <typename T>
__device__ int invert_v(const SolveArgs& a, double* Vbuf, double* Vi, double* pbuf, int* pflag, int* piv,
                        int* flag) {
  const int R = a.R;
  build_v<T>(a, Vbuf);  // V from the Grams, all threads (global loads in parallel)
  bool spd_ok;
  if (R <= 32) {
    if (threadIdx.x < 64) {
      const bool ok = wave_spd_inverse32(Vbuf, R, Vi);
      if (threadIdx.x == 0) pflag[0] = ok ? 1 : 0;
    }
    __syncthreads();
    spd_ok = pflag[0] != 0;
  } else {
    spd_ok = spd_gj_inverse(Vbuf, R, Vi, pbuf, pflag);
  }
  if (spd_ok) return 0;
  return lds_gj_inverse(Vbuf, Vi, R, piv, flag) ? 1 : 2;
}

// V^-1 alone (one block), so that it can run on a second stream while the
// MTTKRP of the same mode is still streaming the tensor.
template <typename T>
__global__ __launch_bounds__(256) void cp_inverse_kernel(SolveArgs a, double* vi_out, int* status_out) {
  __shared__ double Vbuf[kMaxRank * kMaxRank];
  __shared__ double Vi[kMaxRank * kMaxRank];
  __shared__ double pbuf[256];
  __shared__ int flag, piv, pflag[2];
  const int st = invert_v<T>(a, Vbuf, Vi, pbuf, pflag, &piv, &flag);
  __syncthreads();
  for (int e = threadIdx.x; e < a.R * a.R; e += blockDim.x) vi_out[e] = Vi[e];
  if (threadIdx.x == 0) *status_out = st;
}

template <typename T>
__global__ __launch_bounds__(256) void cp_solve_kernel(SolveArgs a) {
  __shared__ double Lbuf[kMaxRank * kMaxRank];
  __shared__ double Vi[kMaxRank * kMaxRank];
  __shared__ double Mr[kSolveRows * kMaxRank];
  __shared__ double pbuf[256];
  __shared__ int flag, piv, ticket, pflag[2];
  const int R = a.R, tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kSolveRows;
  const int nr = (int)min((int64_t)kSolveRows, a.rows - r0);
  // this block's MTTKRP rows, staged once (coalesced)
  for (int e = tid; e < nr * R; e += blockDim.x) Mr[e] = ld<T>(a.mttkrp, r0 * R + e);
  int status = 0;
  if (a.vi_pre) {
    for (int e = tid; e < R * R; e += blockDim.x) Vi[e] = a.vi_pre[e];
    status = *a.vi_status;
    __syncthreads();
  } else {
    status = invert_v<T>(a, Lbuf, Vi, pbuf, pflag, &piv, &flag);
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
  // Sum the per-block partials in block order.  Every block's partials were
  // published before its ticket (release fence above, acquire fence here);
  // this block never read them before, so plain loads see them.  Loads are
  // issued eight blocks at a time so they overlap.
  double mn2 = 0.0;  // ||[[w; A]]||^2 = sum (Hadamard of all Grams) w w^T, for the fused error
  for (int e0 = tid; e0 < R * R; e0 += 4 * 256) {
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = 0; b0 < nblk; b0 += 8) {
      double v[8][4];
#pragma unroll
      for (int bb = 0; bb < 8; ++bb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = e0 + j * 256;
          v[bb][j] = (b0 + bb < nblk && e < R * R) ? a.gram_partial[(int64_t)(b0 + bb) * R * R + e] : 0.0;
        }
#pragma unroll
      for (int bb = 0; bb < 8; ++bb)
#pragma unroll
        for (int j = 0; j < 4; ++j) g[j] += v[bb][j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = e0 + j * 256;
      if (e < R * R) {
        a.gram_out[e] = g[j];
        if (a.errors) {
          double v = g[j];
          for (int m = 0; m < a.ndim; ++m)
            if (m != a.mode) v *= a.grams[m][e];
          if (a.weights) v *= ld<T>(a.weights, e / R) * ld<T>(a.weights, e % R);
          mn2 += v;
        }
      }
    }
  }
  double ip_all = 0.0;
  if (a.ip_partial && tid == 0) {
    for (int b0 = 0; b0 < nblk; b0 += 8) {
      double v[8];
#pragma unroll
      for (int bb = 0; bb < 8; ++bb) v[bb] = (b0 + bb < nblk) ? a.ip_partial[b0 + bb] : 0.0;
#pragma unroll
      for (int bb = 0; bb < 8; ++bb) ip_all += v[bb];
    }
    *a.iprod_out = ip_all;
  }
  if (a.errors) {
    __shared__ double red2[4];
    mn2 = wave_sum(mn2);
    if ((tid & 63) == 0) red2[tid >> 6] = mn2;
    __syncthreads();
    if (tid == 0) {
      const double m2 = (red2[0] + red2[1]) + (red2[2] + red2[3]);
      const double nx2 = *a.normx2;
      a.errors[a.err_slot] = sqrt(fabs(nx2 + m2 - 2.0 * ip_all)) / sqrt(nx2);
    }
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

int tla::cp_solve_launch(int dtype, const double* const* grams, int ndim, int mode, const void* weights,
                         double l2_reg, const void* mttkrp, int64_t rows, int rank, void* factor_out,
                         double* gram_out, double* iprod_out, int* status_out, int* counter, void* ws,
                         size_t ws_bytes, const double* normx2, double* errors, int err_slot,
                         const double* vi_pre, const int* vi_status, void* stream) {
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
  TLA_REQUIRE(!errors || (normx2 && iprod_out && counter), "cp_solve: the fused error needs normx2, iprod, counter");
  a.normx2 = normx2;
  a.errors = errors;
  a.err_slot = err_slot;
  TLA_REQUIRE(!vi_pre || vi_status, "cp_solve: precomputed inverse without status");
  a.vi_pre = vi_pre;
  a.vi_status = vi_status;
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

extern "C" int tla_cp_solve(int dtype, const double* const* grams, int ndim, int mode, const void* weights,
                            double l2_reg, const void* mttkrp, int64_t rows, int rank, void* factor_out,
                            double* gram_out, double* iprod_out, int* status_out, int* counter, void* ws,
                            size_t ws_bytes, void* stream) {
  return tla::cp_solve_launch(dtype, grams, ndim, mode, weights, l2_reg, mttkrp, rows, rank, factor_out, gram_out,
                              iprod_out, status_out, counter, ws, ws_bytes, nullptr, nullptr, 0, nullptr, nullptr,
                              stream);
}

int tla::cp_inverse_launch(int dtype, const double* const* grams, int ndim, int mode, const void* weights,
                           double l2_reg, int rank, double* vi_out, int* status_out, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "cp_inverse: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder && mode >= 0 && mode < ndim, "cp_inverse: order %d mode %d", ndim,
              mode);
  TLA_REQUIRE(rank >= 1 && rank <= kMaxRank, "cp_inverse: rank %d", rank);
  TLA_REQUIRE(vi_out && status_out, "cp_inverse: NULL output");
  SolveArgs a{};
  for (int m = 0; m < ndim; ++m) a.grams[m] = grams[m];
  a.ndim = ndim;
  a.mode = mode;
  a.R = rank;
  a.weights = weights;
  a.l2 = l2_reg;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) cp_inverse_kernel<float><<<1, 256, 0, st>>>(a, vi_out, status_out);
  else cp_inverse_kernel<double><<<1, 256, 0, st>>>(a, vi_out, status_out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
