// FP64 pieces of the ALS sweep: factor Grams, the Hadamard-of-Grams normal
// equations solved by symmetric positive definite elimination (LU with
// partial pivoting as the fallback), the squared norm of the
// tensor and the reconstruction error.  Everything here is R x R or I x R
// sized -- launch/latency bound -- so each step is one small fused kernel and
// the R x R work is done redundantly per workgroup rather than paying a
// second launch.
#include <math.h>

#include "common.h"
#include "inverse.h"

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
// Normal-equations solve.  Each block handles kSolveRows rows of the new
// factor, its partial Gram and partial <M, A>, with V^-1 either computed
// ahead (by the previous solve's finalising block, see next_mode) or rebuilt
// and inverted by every block in registers.
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
  int* status_max;       // nullable: keeps the largest status of every solve since it was zeroed
  int* counter;          // nullable: zero-initialised block counter -> the last block finalises
  double* gram_out;
  double* iprod_out;
  // optional, last mode only: errors[err_slot] = relative error of the model
  // after this update (needs iprod_out), computed by the finalising block
  const double* normx2;
  double* errors;
  int err_slot;
  // optional: V^-1 computed ahead (R x R FP64, global) and its status; a
  // negative status means "not valid any more" (the Grams changed since):
  // the blocks then invert V themselves
  const double* vi_pre;
  const int* vi_status;
  // optional (needs counter): after finalising, the last block also forms
  // and inverts V of mode next_mode from the Grams including this update's,
  // into vi_next / vist_next, for the next solve (-1: none)
  int next_mode;
  double* vi_next;
  int* vist_next;
};

typedef double doublex4 __attribute__((ext_vector_type(4)));

// D = A B + C, A 16 x 4, B 4 x 16 (FP64 MFMA): lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and holds D[l / 16 + 4 e][l % 16], e < 4 (rows
// interleaved by 4, unlike the FP32 16x16x4 form).
__device__ __forceinline__ doublex4 mfma_f64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static_assert(kSolveRows == 64, "the solve's MFMA tiling assigns 16 rows to each of 4 waves");

// RPC: the padded rank (16, 32 or 64) as a compile-time constant, so the
// MFMA k-loops unroll and their LDS operands are requested ahead.
template <typename T, int RPC>
__global__ __launch_bounds__(256) void cp_solve_kernel(SolveArgs a) {
  __shared__ double Lbuf[kMaxRank * kMaxRank];  // V, then the rounded new rows (kSolveRows x RP)
  __shared__ double Vi[kMaxRank * kMaxRank];    // V^-1, RP x RP
  __shared__ double Mr[kSolveRows * kMaxRank];  // this block's MTTKRP rows, kSolveRows x RP, zero padded
  __shared__ double pbuf[6 * kMaxRank];
  __shared__ int flag, piv, ticket, pflag[2];
  constexpr int RP = RPC, RT = RPC / 16;
  const int R = a.R, tid = threadIdx.x;
  constexpr int wdt = sizeof(T) == 8 ? TLA_F64 : TLA_F32;
  const int lane = tid & 63, w = tid >> 6, lr = lane & 15, lk = lane >> 4;
  const int64_t r0 = (int64_t)blockIdx.x * kSolveRows;
  const int nr = (int)min((int64_t)kSolveRows, a.rows - r0);
  // Every global load of the prologue in one batch: the MTTKRP rows, the
  // precomputed V^-1 (copied whether or not its status says it is valid;
  // an invalid one is overwritten below) and that status.
  const int pre_status = a.vi_pre ? *a.vi_status : -1;  // the same word for every thread
#pragma unroll
  for (int e0 = 0; e0 < kSolveRows * RP; e0 += 256) {
    const int e = e0 + tid, i = e / RP, c = e % RP;
    Mr[e] = (i < nr && c < R) ? ld<T>(a.mttkrp, (r0 + i) * R + c) : 0.0;
  }
  if (a.vi_pre) {
#pragma unroll
    for (int e0 = 0; e0 < RP * RP; e0 += 256) {
      const int e = e0 + tid;
      if (e < RP * RP) {
        const int r = e / RP, s = e % RP;
        Vi[e] = (r < R && s < R) ? a.vi_pre[r * R + s] : 0.0;
      }
    }
  }
  int status = pre_status;
  if (status < 0) {
    __syncthreads();  // the speculative copy is done before invert_v reuses Vi
    status = invert_v<256>(a, a.ndim, a.mode, -1, nullptr, a.weights, wdt, a.l2, R, Lbuf, Vi, pbuf, pflag, &piv, &flag);
  }
  if (blockIdx.x == 0 && tid == 0) {
    if (a.status) *a.status = status;
    if (a.status_max && status > *a.status_max) *a.status_max = status;
  }
  __syncthreads();
  // New rows A = M V^-1 (V symmetric) on FP64 MFMA: wave w computes rows
  // 16 w .. 16 w + 15, every 16-column tile.  The rounded values go to the
  // factor, to LDS (reusing Lbuf) for the Gram, and into <M, A>.
  double* An = Lbuf;
  double ip = 0.0;
  {
    doublex4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < RP; k0 += 4) {
      const double av = Mr[(16 * w + lr) * RP + k0 + lk];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t < RT) acc[t] = mfma_f64(av, Vi[(k0 + lk) * RP + 16 * t + lr], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t < RT) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * w + lk + 4 * e, col = 16 * t + lr;
          const bool ok = row < nr && col < R;
          const T out = (T)acc[t][e];
          if (ok) ((T*)a.factor)[(r0 + row) * R + col] = out;
          const double o = ok ? (double)out : 0.0;
          An[row * RP + col] = o;
          ip = fma(Mr[row * RP + col], o, ip);
        }
      }
    }
  }
  __syncthreads();
  // Partial Gram A^T A of these rows (zero rows past nr), one 16 x 16 tile
  // per wave and step
  for (int tile = w; tile < RT * RT; tile += 4) {
    const int tr = tile / RT, tc = tile % RT;
    doublex4 g = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < kSolveRows; k0 += 4)
      g = mfma_f64(An[(k0 + lk) * RP + 16 * tr + lr], An[(k0 + lk) * RP + 16 * tc + lr], g);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = 16 * tr + lk + 4 * e, s = 16 * tc + lr;
      if (r < R && s < R) a.gram_partial[(int64_t)blockIdx.x * R * R + r * R + s] = g[e];
    }
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
  if (a.next_mode >= 0) {
    // V^-1 for the next solve, from the Grams with this update's (gram_out,
    // written above by this block; the barrier orders it for every thread)
    __syncthreads();
    const int st = invert_v<256>(a, a.ndim, a.next_mode, a.mode, a.gram_out, a.weights, wdt, a.l2, R, Lbuf, Vi, pbuf,
                                 pflag, &piv, &flag);
    for (int e = tid; e < R * R; e += blockDim.x) a.vi_next[e] = Vi[(e / R) * RP + e % R];
    if (tid == 0) *a.vist_next = st;
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


// ---------------------------------------------------------------------------
// Multiplicative update of a non-negative CP factor
// (tensorly/decomposition/_nn_cp.py:111-135):
//   accum = (Hadamard of the other modes' Grams) * w w^T
//   A <- A * max(M, eps) / max(A accum, eps)
// accum is formed in FP64 in LDS; one wave per row, lane r owns column r and
// reads the whole row before the wave writes it back (in place).
// ---------------------------------------------------------------------------
struct MuArgs {
  const double* grams[kMaxOrder];
  int ndim, mode, R;
  const void* weights;
  int wdt;
  const void* M;
  void* A;
  int64_t rows;
  double eps;
  void* out;  // op 1
};

// op 0: the multiplicative update above (in place).
// op 1: out = A accum - M, the gradient of 0.5 ||X - [[w; A]]||^2 with
//       respect to the mode's factor without forming the reconstruction
//       (MTTKRP of the model = A accum).
template <typename T, int OP>
__global__ __launch_bounds__(256) void nn_mu_update_kernel(MuArgs a) {
  __shared__ double acc[kMaxRank * kMaxRank];
  const int R = a.R;
  build_v(a, a.ndim, a.mode, -1, nullptr, a.weights, a.wdt, 0.0, R, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T* A = (T*)a.A;
  const T* M = (const T*)a.M;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave; i < a.rows; i += nw) {
    const T* row = A + i * R;
    double den = 0.0;
    for (int k = 0; k < R; ++k) den = fma((double)row[k], lane < R ? acc[k * R + lane] : 0.0, den);
    if (lane < R) {
      if constexpr (OP == 0) {
        const double num = fmax((double)M[i * R + lane], a.eps);
        const double aval = (double)row[lane];
        A[i * R + lane] = (T)(aval * num / fmax(den, a.eps));
      } else {
        ((T*)a.out)[i * R + lane] = (T)(den - (double)M[i * R + lane]);
      }
    }
  }
}


// ---------------------------------------------------------------------------
// HALS non-negative least squares for one CP factor (tensorly/solvers/
// nnls.py:5-173 as called by non_negative_parafac_hals, _nn_cp.py:300-311):
// UtU = (Hadamard of the other modes' Grams) * w w^T, UtM = M^T, V = A^T.
// Each iteration updates the rows k = 0..R-1 of V in order (Gauss-Seidel),
//   V[k,:] <- max((UtM[k,:] - UtU[k,:] V + UtU[k,k] V[k,:] - s) / UtU[k,k], 0)
// and stops when the iteration's error sum is below tol times the first
// iteration's; as in the reference (nnls.py:147, ||V - newV||^2 with the new
// row broadcast against all of V) each row update adds
// sum_{j < R} (V[j, i] - newV[i])^2 over the columns i.  The columns of V (rows of A) are independent within an
// iteration: one workgroup, each thread keeps its rows' values in registers
// for the R updates; the squared update is summed in a fixed order.
// ---------------------------------------------------------------------------
struct HalsArgs {
  const double* grams[kMaxOrder];
  int ndim, mode, R;
  const void* weights;
  int wdt;
  const void* M;
  void* A;
  int64_t rows;
  double sparsity;
  int has_sparsity;
  int n_iter_max;
  double tol;
  int* iters_out;
};

constexpr int kHalsThreads = 512;

template <typename T, int RT>
__global__ __launch_bounds__(kHalsThreads) void hals_nnls_kernel(HalsArgs a) {
  __shared__ double utu[RT * RT];
  __shared__ double acc[kMaxRank * kMaxRank];
  __shared__ double red[kHalsThreads / 64];
  __shared__ int stop;
  const int R = a.R, tid = threadIdx.x;
  build_v(a, a.ndim, a.mode, -1, nullptr, a.weights, a.wdt, 0.0, R, acc);
  for (int e = tid; e < RT * RT; e += blockDim.x) {
    const int k = e / RT, l = e % RT;
    utu[e] = (k < R && l < R) ? acc[k * R + l] : 0.0;
  }
  __syncthreads();
  T* A = (T*)a.A;
  const T* M = (const T*)a.M;
  double err0 = 0.0;
  int it = 0;
  for (; it < a.n_iter_max; ++it) {
    double err = 0.0;
    for (int64_t i = tid; i < a.rows; i += blockDim.x) {
      // UtU is loop-invariant; an opaque copy of its address per row keeps
      // the compiler from holding all RT * RT values in registers
      const double* U = utu;
      asm volatile("" : "+v"(U));
      // at RT = 64 the MTTKRP row is read per update (v and m together would
      // need every register)
      constexpr bool kMRegs = RT <= 48;
      double v[RT], m[kMRegs ? RT : 1];
#pragma unroll
      for (int l = 0; l < RT; ++l) {
        v[l] = l < R ? (double)A[i * R + l] : 0.0;
        if constexpr (kMRegs) m[l] = l < R ? (double)M[i * R + l] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < RT; ++k) {
        const double d = U[k * RT + k];
        if (d != 0.0) {
          double dot = 0.0;
#pragma unroll
          for (int l = 0; l < RT; ++l) dot = fma(U[k * RT + l], v[l], dot);
          double mk;
          if constexpr (kMRegs) mk = m[k];
          else mk = (double)M[i * R + k];  // d != 0 only for k < R
          double num = mk - dot + d * v[k];
          if (a.has_sparsity) num -= a.sparsity;
          const double nv = fmax(num / d, 0.0);
          // the stopping sum subtracts the new value from every entry of
          // this column of V (the reference's broadcast ||V - newV||^2)
#pragma unroll
          for (int l = 0; l < RT; ++l)
            if (l < R) {
              const double df = v[l] - nv;
              err = fma(df, df, err);
            }
          v[k] = nv;
        }
      }
#pragma unroll
      for (int l = 0; l < RT; ++l)
        if (l < R) A[i * R + l] = (T)v[l];
    }
    err = wave_sum(err);
    if ((tid & 63) == 0) red[tid >> 6] = err;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
      if (it == 0) red[0] = s, err0 = s;
      stop = (s < a.tol * (it == 0 ? s : err0)) ? 1 : 0;
    }
    __syncthreads();
    const bool done = stop != 0;
    __syncthreads();
    if (done) {
      ++it;
      break;
    }
  }
  if (tid == 0 && a.iters_out) *a.iters_out = it;
}


// Multi-workgroup HALS: one launch per iteration (rows over the grid, the
// workgroup's squared update summed in a fixed order into partial[block]),
// then one single-thread launch that sums the partials in block order and
// sets the stop flag; later iteration launches see the flag and return.
// state: [0] stop flag, [1] iterations run; err0 at state_d[0].
template <int RT>
__global__ void hals_prep_kernel(HalsArgs a, double* utu_g, int* state, double* err0) {
  __shared__ double acc[kMaxRank * kMaxRank];
  build_v(a, a.ndim, a.mode, -1, nullptr, a.weights, a.wdt, 0.0, a.R, acc);
  for (int e = threadIdx.x; e < RT * RT; e += blockDim.x) {
    const int k = e / RT, l = e % RT;
    utu_g[e] = (k < a.R && l < a.R) ? acc[k * a.R + l] : 0.0;
  }
  if (threadIdx.x == 0) {
    state[0] = 0;
    state[1] = 0;
    err0[0] = 0.0;
  }
}

constexpr int kHalsRowsPerBlock = 64;

template <typename T, int RT>
__global__ __launch_bounds__(kHalsRowsPerBlock) void hals_iter_kernel(HalsArgs a, const double* utu_g,
                                                                      double* partial, const int* state) {
  if (state[0]) return;
  __shared__ double utu[RT * RT];
  const int tid = threadIdx.x;
  for (int e = tid; e < RT * RT; e += blockDim.x) utu[e] = utu_g[e];
  __syncthreads();
  const int R = a.R;
  T* A = (T*)a.A;
  const T* M = (const T*)a.M;
  double err = 0.0;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  if (i < a.rows) {
    const double* U = utu;
    asm volatile("" : "+v"(U));
    constexpr bool kMRegs = RT <= 48;
    double v[RT], m[kMRegs ? RT : 1];
#pragma unroll
    for (int l = 0; l < RT; ++l) {
      v[l] = l < R ? (double)A[i * R + l] : 0.0;
      if constexpr (kMRegs) m[l] = l < R ? (double)M[i * R + l] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < RT; ++k) {
      const double d = U[k * RT + k];
      if (d != 0.0) {
        double dot = 0.0;
#pragma unroll
        for (int l = 0; l < RT; ++l) dot = fma(U[k * RT + l], v[l], dot);
        double mk;
        if constexpr (kMRegs) mk = m[k];
        else mk = (double)M[i * R + k];
        double num = mk - dot + d * v[k];
        if (a.has_sparsity) num -= a.sparsity;
        const double nv = fmax(num / d, 0.0);
        // the stopping sum subtracts the new value from every entry of
        // this column of V (the reference's broadcast ||V - newV||^2)
#pragma unroll
        for (int l = 0; l < RT; ++l)
          if (l < R) {
            const double df = v[l] - nv;
            err = fma(df, df, err);
          }
        v[k] = nv;
      }
    }
#pragma unroll
    for (int l = 0; l < RT; ++l)
      if (l < R) A[i * R + l] = (T)v[l];
  }
  err = wave_sum(err);
  if (tid == 0) partial[blockIdx.x] = err;
}

__global__ void hals_decide_kernel(const double* partial, int nblocks, double tol, int* state, double* err0) {
  if (threadIdx.x != 0 || state[0]) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += partial[b];
  if (state[1] == 0) err0[0] = s;
  state[1] += 1;
  if (s < tol * err0[0]) state[0] = 1;
}

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
                         const double* vi_pre, const int* vi_status, int next_mode, double* vi_next,
                         int* vist_next, int* status_max, void* stream) {
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
  a.status_max = status_max;
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
  TLA_REQUIRE(next_mode < ndim && (next_mode < 0 || (counter && vi_next && vist_next)),
              "cp_solve: the next mode's inverse needs counter, vi_next and vist_next");
  a.next_mode = next_mode < 0 ? -1 : next_mode;
  a.vi_next = vi_next;
  a.vist_next = vist_next;
  hipStream_t st = (hipStream_t)stream;
  const int rp = solve_pad(rank);
#define TLA_SOLVE_CASE(RPV)                                                 \
  if (rp == RPV) {                                                          \
    if (dtype == TLA_F32) cp_solve_kernel<float, RPV><<<nblk, 256, 0, st>>>(a);  \
    else cp_solve_kernel<double, RPV><<<nblk, 256, 0, st>>>(a);               \
  }
  TLA_SOLVE_CASE(16)
  TLA_SOLVE_CASE(32)
  TLA_SOLVE_CASE(64)
#undef TLA_SOLVE_CASE
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
                              -1, nullptr, nullptr, nullptr, stream);
}

extern "C" int tla_nn_mu_update(int dtype, void* factor, const void* mttkrp, int64_t rows, int rank,
                                const double* const* grams, int ndim, int mode, const void* weights, int wdtype,
                                double eps, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "nn_mu_update: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "nn_mu_update: order %d", ndim);
  TLA_REQUIRE(mode >= 0 && mode < ndim, "nn_mu_update: mode %d", mode);
  TLA_REQUIRE(rank >= 1 && rank <= kMaxRank, "nn_mu_update: rank %d not in [1, %d]", rank, kMaxRank);
  TLA_REQUIRE(rows >= 0, "nn_mu_update: rows = %lld", (long long)rows);
  TLA_REQUIRE(factor && mttkrp, "nn_mu_update: NULL pointer");
  TLA_REQUIRE(!weights || wdtype == TLA_F32 || wdtype == TLA_F64, "nn_mu_update: bad weight dtype %d", wdtype);
  if (rows == 0) return TLA_OK;
  MuArgs a{};
  for (int m = 0; m < ndim; ++m) {
    TLA_REQUIRE(m == mode || grams[m] != nullptr, "nn_mu_update: gram %d is NULL", m);
    a.grams[m] = grams[m];
  }
  a.ndim = ndim; a.mode = mode; a.R = rank;
  a.weights = weights; a.wdt = wdtype;
  a.M = mttkrp; a.A = factor; a.rows = rows; a.eps = eps;
  const int64_t blocks = (rows + 3) / 4 < 1024 ? (rows + 3) / 4 : 1024;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) nn_mu_update_kernel<float, 0><<<(unsigned)blocks, 256, 0, st>>>(a);
  else nn_mu_update_kernel<double, 0><<<(unsigned)blocks, 256, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" int tla_cp_lstsq_grad_mode(int dtype, const void* factor, const void* mttkrp, void* grad, int64_t rows,
                                      int rank, const double* const* grams, int ndim, int mode, const void* weights,
                                      int wdtype, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "cp_lstsq_grad: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "cp_lstsq_grad: order %d", ndim);
  TLA_REQUIRE(mode >= 0 && mode < ndim, "cp_lstsq_grad: mode %d", mode);
  TLA_REQUIRE(rank >= 1 && rank <= kMaxRank, "cp_lstsq_grad: rank %d not in [1, %d]", rank, kMaxRank);
  TLA_REQUIRE(rows >= 0, "cp_lstsq_grad: rows = %lld", (long long)rows);
  TLA_REQUIRE(factor && mttkrp && grad, "cp_lstsq_grad: NULL pointer");
  TLA_REQUIRE(grad != factor && grad != mttkrp, "cp_lstsq_grad: grad must not alias its inputs");
  TLA_REQUIRE(!weights || wdtype == TLA_F32 || wdtype == TLA_F64, "cp_lstsq_grad: bad weight dtype %d", wdtype);
  if (rows == 0) return TLA_OK;
  MuArgs a{};
  for (int m = 0; m < ndim; ++m) {
    TLA_REQUIRE(m == mode || grams[m] != nullptr, "cp_lstsq_grad: gram %d is NULL", m);
    a.grams[m] = grams[m];
  }
  a.ndim = ndim; a.mode = mode; a.R = rank;
  a.weights = weights; a.wdt = wdtype;
  a.M = mttkrp; a.A = const_cast<void*>(factor); a.rows = rows; a.eps = 0.0; a.out = grad;
  const int64_t blocks = (rows + 3) / 4 < 1024 ? (rows + 3) / 4 : 1024;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) nn_mu_update_kernel<float, 1><<<(unsigned)blocks, 256, 0, st>>>(a);
  else nn_mu_update_kernel<double, 1><<<(unsigned)blocks, 256, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

static size_t hals_align(size_t b) { return (b + 255) & ~size_t(255); }

// Iteration counts up to this run as one launch per iteration over the grid;
// longer runs (exact mode) stay in one workgroup.
constexpr int kHalsMaxLaunches = 1000;

extern "C" size_t tla_hals_nnls_workspace_size(int64_t rows, int rank) {
  (void)rank;
  const int64_t nb = (rows + kHalsRowsPerBlock - 1) / kHalsRowsPerBlock;
  return hals_align(64 * 64 * 8) + hals_align(size_t(nb > 0 ? nb : 1) * 8) + hals_align(16) + hals_align(8);
}

extern "C" int tla_hals_nnls(int dtype, void* factor, const void* mttkrp, int64_t rows, int rank,
                             const double* const* grams, int ndim, int mode, const void* weights, int wdtype,
                             double sparsity, int has_sparsity, int n_iter_max, double tol, int* iters_out,
                             void* workspace, size_t workspace_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "hals_nnls: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "hals_nnls: order %d", ndim);
  TLA_REQUIRE(mode >= 0 && mode < ndim, "hals_nnls: mode %d", mode);
  TLA_REQUIRE(rank >= 1 && rank <= kMaxRank, "hals_nnls: rank %d not in [1, %d]", rank, kMaxRank);
  TLA_REQUIRE(rows >= 0, "hals_nnls: rows = %lld", (long long)rows);
  TLA_REQUIRE(n_iter_max >= 0, "hals_nnls: n_iter_max = %d", n_iter_max);
  TLA_REQUIRE(factor && mttkrp, "hals_nnls: NULL pointer");
  TLA_REQUIRE(!weights || wdtype == TLA_F32 || wdtype == TLA_F64, "hals_nnls: bad weight dtype %d", wdtype);
  if (rows == 0 || n_iter_max == 0) return TLA_OK;
  HalsArgs a{};
  for (int m = 0; m < ndim; ++m) {
    TLA_REQUIRE(m == mode || grams[m] != nullptr, "hals_nnls: gram %d is NULL", m);
    a.grams[m] = grams[m];
  }
  a.ndim = ndim; a.mode = mode; a.R = rank;
  a.weights = weights; a.wdt = wdtype;
  a.M = mttkrp; a.A = factor; a.rows = rows;
  a.sparsity = sparsity; a.has_sparsity = has_sparsity;
  a.n_iter_max = n_iter_max; a.tol = tol; a.iters_out = iters_out;
  hipStream_t st = (hipStream_t)stream;
  const int rt = rank <= 16 ? 16 : rank <= 32 ? 32 : rank <= 48 ? 48 : 64;
  if (workspace && n_iter_max <= kHalsMaxLaunches) {
    TLA_REQUIRE(workspace_bytes >= tla_hals_nnls_workspace_size(rows, rank), "hals_nnls: workspace too small");
    const int64_t nb = (rows + kHalsRowsPerBlock - 1) / kHalsRowsPerBlock;
    TLA_REQUIRE(nb <= (int64_t)1 << 30, "hals_nnls: %lld rows", (long long)rows);
    char* w = (char*)workspace;
    double* utu_g = (double*)w;
    double* partial = (double*)(w + hals_align(64 * 64 * 8));
    int* state = (int*)(w + hals_align(64 * 64 * 8) + hals_align(size_t(nb) * 8));
    double* err0 = (double*)((char*)state + hals_align(16));
#define TLA_HALS_MULTI(T_, RT_)                                                                       \
  do {                                                                                                \
    hals_prep_kernel<RT_><<<1, 256, 0, st>>>(a, utu_g, state, err0);                                  \
    for (int it = 0; it < n_iter_max; ++it) {                                                         \
      hals_iter_kernel<T_, RT_><<<(unsigned)nb, kHalsRowsPerBlock, 0, st>>>(a, utu_g, partial, state); \
      hals_decide_kernel<<<1, 64, 0, st>>>(partial, (int)nb, tol, state, err0);                       \
    }                                                                                                 \
  } while (0)
#define TLA_HALS_RT(T_)                                \
  switch (rt) {                                        \
    case 16: TLA_HALS_MULTI(T_, 16); break;            \
    case 32: TLA_HALS_MULTI(T_, 32); break;            \
    case 48: TLA_HALS_MULTI(T_, 48); break;            \
    default: TLA_HALS_MULTI(T_, 64); break;            \
  }
    if (dtype == TLA_F32) {
      TLA_HALS_RT(float)
    } else {
      TLA_HALS_RT(double)
    }
#undef TLA_HALS_RT
#undef TLA_HALS_MULTI
    TLA_LAUNCH_CHECK();
    if (iters_out) {
      TLA_HIP(hipMemcpyAsync(iters_out, state + 1, sizeof(int), hipMemcpyDeviceToDevice, st));
    }
    return TLA_OK;
  }
#define TLA_HALS(T_)                                                                  \
  switch (rt) {                                                                       \
    case 16: hals_nnls_kernel<T_, 16><<<1, kHalsThreads, 0, st>>>(a); break;          \
    case 32: hals_nnls_kernel<T_, 32><<<1, kHalsThreads, 0, st>>>(a); break;          \
    case 48: hals_nnls_kernel<T_, 48><<<1, kHalsThreads, 0, st>>>(a); break;          \
    default: hals_nnls_kernel<T_, 64><<<1, kHalsThreads, 0, st>>>(a); break;          \
  }
  if (dtype == TLA_F32) {
    TLA_HALS(float)
  } else {
    TLA_HALS(double)
  }
#undef TLA_HALS
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
