// Inverse of the CP-ALS normal-equations matrix
//   V = (Hadamard product of the other modes' FP64 Grams + l2 I) * w w^T
// (tensorly/decomposition/_cp.py:411-427 solves with it), for R <= 64, by
// one workgroup.  Used by the solve kernel (inline, or for the next mode
// from its finalising block) and by the persistent MTTKRP kernels, where one
// extra workgroup inverts V of the mode whose solve follows while the others
// stream the tensor (InvJob).
#pragma once
#include "common.h"

namespace tla {

// The R x R systems are padded to RP = 16, 32 or 64.
__host__ __device__ constexpr int solve_pad(int R) { return R <= 16 ? 16 : (R <= 32 ? 32 : 64); }

// LDS bytes run_inv_job needs (V, padded V^-1, pivot buffers, flags).
__host__ __device__ constexpr size_t inv_job_lds_bytes(int R) {
  return size_t(8) * (size_t(R) * R + size_t(solve_pad(R)) * solve_pad(R) + 6 * solve_pad(R)) + 16;
}

// 1 / p to full double precision: the hardware reciprocal refined by two
// Newton steps (each doubles the correct bits); p > 0 and finite here.
__device__ __forceinline__ double recip_f64(double p) {
  double r = __builtin_amdgcn_rcp(p);
  r = fma(r, fma(-p, r, 1.0), r);
  r = fma(r, fma(-p, r, 1.0), r);
  return r;
}

// V^-1 of the symmetric positive definite V (R x R, stride R, in LDS, left
// intact) by Gauss-Jordan without pivoting on [V_p | I], V_p = diag(V, I) of
// size RP.  The NT threads of the block hold [V_p | I] in registers: TPR =
// NT / RP consecutive threads own one row, CPT = 2 RP / TPR columns each.
// Per pivot k the owners of row k publish it unscaled and the owners of
// column k publish that column through LDS (double-buffered: one barrier
// per pivot); every thread then forms its multiplier x_ik / x_kk and updates
// its entries.  For SPD V the pivots are the positive Schur-complement
// diagonals; if one is not positive and finite the function returns false
// (the caller falls back to partial pivoting).  Writes Vi = diag(V^-1, I)
// (RP x RP, stride RP).  pbuf: 6 RP doubles.
template <int RP, int NT>
__device__ __forceinline__ bool block_gj_inverse(const double* V, int R, double* Vi, double* pbuf, int* pflag) {
  constexpr int TPR = NT / RP;
  constexpr int CPT = 2 * RP / TPR;
  static_assert(TPR >= 2 && TPR <= 64 && CPT >= 1, "thread layout");
  const int tid = threadIdx.x;
  const int i = tid / TPR, seg = tid % TPR, c0 = seg * CPT;
  double x[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = c0 + j;
    double v;
    if (c < RP) v = (i < R && c < R) ? V[i * R + c] : (i == c ? 1.0 : 0.0);
    else v = (c - RP == i) ? 1.0 : 0.0;
    x[j] = v;
  }
  if (tid == 0) *pflag = 0;
  bool bad = false;
  for (int kb = 0; kb < TPR / 2; ++kb) {
#pragma unroll
    for (int kk = 0; kk < CPT; ++kk) {
      const int k = kb * CPT + kk;
      double* row = pbuf + (k & 1) * (3 * RP);
      double* col = row + 2 * RP;
      if (i == k) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) row[c0 + j] = x[j];
      }
      if (seg == kb) col[i] = x[kk];
      __syncthreads();
      const double p = col[k];
      const double inv = recip_f64(p);
      if (i == k) {
        bad = bad || !(p > 0.0) || !isfinite(p);
#pragma unroll
        for (int j = 0; j < CPT; ++j) x[j] *= inv;
      } else {
        // the multiplier by a true division: a row equal to the pivot row
        // gets exactly 1 and cancels to exact zeros, so a V with two equal
        // rows (duplicated factor columns) meets an exactly zero pivot and
        // is reported singular, as LAPACK's LU finds it
        const double f = col[i] / p;
#pragma unroll
        for (int j = 0; j < CPT; ++j) x[j] = fma(-f, row[c0 + j], x[j]);
      }
    }
  }
  if (bad) *pflag = 1;
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = c0 + j;
    if (c >= RP) Vi[i * RP + (c - RP)] = x[j];
  }
  __syncthreads();
  return *pflag == 0;
}

// Gauss-Jordan inverse with partial pivoting (fallback for V that is not
// numerically positive definite).  A is destroyed; Vi is R x R, stride R.
// Returns false if singular: a pivot column with no nonzero entry, the test
// of LAPACK's getrf.  Multipliers are formed against the unscaled pivot row
// (A[i][k] / A[k][k]), so rows equal to the pivot row cancel exactly.
__device__ inline bool lds_gj_inverse(double* A, double* Vi, int R, int* piv_shared, int* flag) {
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
    const double akk = A[k * R + k];
    for (int e = tid; e < R * R; e += blockDim.x) {
      const int i = e / R, j = e % R;
      if (i != k) {
        const double f = A[i * R + k] / akk;
        if (j != k) A[i * R + j] -= f * A[k * R + j];
        Vi[i * R + j] -= f * Vi[k * R + j];
      }
    }
    __syncthreads();
    for (int i = tid; i < R; i += blockDim.x)
      if (i != k) A[i * R + k] = 0.0;
    for (int j = tid; j < R; j += blockDim.x) { A[k * R + j] /= akk; Vi[k * R + j] /= akk; }
    __syncthreads();
  }
  return true;
}

__device__ __forceinline__ double weight_at(const void* w, int wdt, int i) {
  return wdt == TLA_F64 ? ((const double*)w)[i] : (double)((const float*)w)[i];
}

// V of mode `mode` (compact R x R into LDS): Hadamard product of the other
// modes' Grams ga.grams[m] (the Gram of mode upd_mode read from g_upd
// instead), + l2 on the diagonal, times w w^T (w nullable, dtype wdt).
template <typename GA>
__device__ __forceinline__ void build_v(const GA& ga, int ndim, int mode, int upd_mode, const double* g_upd, const void* w, int wdt,
                        double l2, int R, double* V) {
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    const int r = e / R, s = e % R;
    double v = 1.0;
    for (int m = 0; m < ndim; ++m)
      if (m != mode) v *= (m == upd_mode ? g_upd : ga.grams[m])[e];
    if (r == s) v += l2;
    if (w) v *= weight_at(w, wdt, r) * weight_at(w, wdt, s);
    V[e] = v;
  }
  __syncthreads();
}

// V (built into Vbuf, R x R) and V^-1 into Vi (RP x RP, stride RP; outside
// the leading R x R block it holds identity or zeros).  Returns 0 (SPD
// elimination), 1 (partial-pivoting fallback) or 2 (singular).  NT threads.
template <int NT, typename GA>
__device__ __forceinline__ int invert_v(const GA& ga, int ndim, int mode, int upd_mode, const double* g_upd, const void* w, int wdt,
                        double l2, int R, double* Vbuf, double* Vi, double* pbuf, int* pflag, int* piv, int* flag) {
  const int RP = solve_pad(R), tid = threadIdx.x;
  build_v(ga, ndim, mode, upd_mode, g_upd, w, wdt, l2, R, Vbuf);
  bool ok;
  if (RP == 16) ok = block_gj_inverse<16, NT>(Vbuf, R, Vi, pbuf, pflag);
  else if (RP == 32) ok = block_gj_inverse<32, NT>(Vbuf, R, Vi, pbuf, pflag);
  else ok = block_gj_inverse<64, NT>(Vbuf, R, Vi, pbuf, pflag);
  if (ok) return 0;
  // partial pivoting on the compact V (intact: the elimination above ran in
  // registers), then spread the compact inverse to stride RP
  const bool fine = lds_gj_inverse(Vbuf, Vi, R, piv, flag);
  for (int e = tid; e < RP * RP; e += blockDim.x) {
    const int r = e / RP, s = e % RP;
    Vbuf[e] = (r < R && s < R) ? Vi[r * R + s] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < RP * RP; e += blockDim.x) Vi[e] = Vbuf[e];
  __syncthreads();
  return fine ? 1 : 2;
}

// Grams stored back to back (mode m at base + m * stride), as build_v's
// `ga.grams[m]`.
struct StridedGrams {
  const double* base;
  int64_t stride;
  __device__ const double* operator[](int m) const { return base + m * stride; }
};
struct StridedGramArgs {
  StridedGrams grams;
};

// One workgroup of NT threads: V^-1 of `mode` (compact R x R FP64) into vi
// and the status (0 / 1 / 2 as invert_v) into *vist; the Grams are R x R
// back to back at `grams`; lds holds inv_job_lds_bytes(R) bytes.  Scalar
// arguments only, so callers keep their kernel arguments in place.
template <int NT>
__device__ __noinline__ void run_inv_job(const double* grams, const void* w, int wdt, double l2, int ndim, int mode,
                                         int R, double* vi, int* vist, char* lds) {
  const int RP = solve_pad(R);
  double* V = (double*)lds;
  double* Vi = V + R * R;
  double* pb = Vi + RP * RP;
  int* ints = (int*)(pb + 6 * RP);
  const StridedGramArgs ga{{grams, (int64_t)R * R}};
  const int st = invert_v<NT>(ga, ndim, mode, -1, nullptr, w, wdt, l2, R, V, Vi, pb, ints, ints + 2, ints + 3);
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) vi[e] = Vi[(e / R) * RP + e % R];
  if (threadIdx.x == 0) *vist = st;
}

// In a kernel whose job descriptor is `job`: true (and the job done) for the
// workgroup that runs it, the last one of the grid.
template <int NT>
__device__ __forceinline__ bool take_inv_job(const InvJob& job, char* lds) {
  if (job.R == 0 || blockIdx.x != gridDim.x - 1) return false;
  run_inv_job<NT>(job.grams, job.weights, job.wdt, job.l2, job.ndim, job.mode, job.R, job.vi, job.vist, lds);
  return true;
}

}  // namespace tla
