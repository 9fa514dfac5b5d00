// AO-ADMM factor update of constrained CP (tensorly/decomposition/
// _constrained_cp.py:166-424) and the proximal operators it applies
// (tensorly/tenalg/proximal.py), for one factor of at most 64 columns.
//
// One workgroup runs the whole inner ADMM loop of tensorly/solvers/admm.py
// (rho = trace(UtU) / R; per iteration x_split = (UtU + rho I)^-1 (UtM +
// rho (x + dual))^T, x = prox(x_split^T - dual), dual += x - x_split^T, stop
// when ||x - x_split^T|| < tol ||x|| and ||x - x_old|| < tol ||dual||).  The
// working arrays are FP64 rows x R matrices in the workspace (row-major like
// the factor); (UtU + rho I)^-1 is formed once per call and kept in LDS.
//
// Proximal operators (rows x R matrix v -> x, FP64):
//   elementwise / whole-matrix  non_negative, l1_reg, l2_reg, l2_square_reg,
//                               normalize, hard_sparsity, normalized_sparsity
//                               (block reductions; the l0 selection is a
//                               radix select on the bits of |v|)
//   per column                  simplex, soft_sparsity (projection threshold
//                               by Michelot's fixed point, one wave per
//                               column), smoothness (tridiagonal LU, one
//                               thread per column), monotonicity and
//                               unimodality (the reference's max-of-means
//                               construction, O(rows^2) per column).
#include <math.h>

#include "common.h"
#include "inverse.h"

namespace tla {

constexpr int kAdmmThreads = 512;
constexpr int kAdmmWaves = kAdmmThreads / 64;

struct ProxBufs {
  const double* v;  // input, rows x R
  double* x;        // output, rows x R
  double* t1;       // scratch, rows x R each
  double* t2;
  double* t3;
  int64_t rows;
  int R;
  int kind;
  double param;
};

struct ProxShared {
  double red[kAdmmWaves];
  int hist[256];
  unsigned long long prefix;
  long long remaining;
  long long cut;
  int jstar[kMaxRank];
};

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum / max over the workgroup, the same value returned to every thread
// (wave partials combined in wave order: repeatable).
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  return s;
}

__device__ double block_max(double v, double* red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = fmax(s, red[w]);
  return s;
}

__device__ __forceinline__ unsigned long long abs_key(double v) {
  return (unsigned long long)__double_as_longlong(fabs(v));
}

// Keep the entries whose rank in decreasing |v| (flat row-major order, equal
// magnitudes ranked later index first, as a stable double argsort does) is
// below k; zero the rest (tensorly/tenalg/proximal.py:634-656).
__device__ void hard_threshold_block(const double* v, double* x, int64_t n, double k, ProxShared& sh) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int64_t kk = k <= 0.0 ? 0 : (int64_t)ceil(k);
  if (kk >= n) {
    for (int64_t e = tid; e < n; e += nt) x[e] = v[e];
    return;
  }
  if (kk == 0) {
    for (int64_t e = tid; e < n; e += nt) x[e] = 0.0;
    return;
  }
  // MSB-first radix select of the kk-th largest key, 8 bits per pass
  if (tid == 0) {
    sh.prefix = 0;
    sh.remaining = kk;
  }
  for (int shift = 56; shift >= 0; shift -= 8) {
    for (int i = tid; i < 256; i += nt) sh.hist[i] = 0;
    __syncthreads();
    const unsigned long long prefix = sh.prefix;
    const unsigned long long hi = shift == 56 ? 0ull : (~0ull << (shift + 8));
    for (int64_t e = tid; e < n; e += nt) {
      const unsigned long long key = abs_key(v[e]);
      if ((key & hi) == (prefix & hi)) atomicAdd(&sh.hist[(key >> shift) & 255], 1);
    }
    __syncthreads();
    if (tid == 0) {
      long long acc = 0;
      int d = 255;
      for (; d > 0; --d) {
        if (acc + sh.hist[d] >= sh.remaining) break;
        acc += sh.hist[d];
      }
      sh.prefix = prefix | ((unsigned long long)d << shift);
      sh.remaining -= acc;
    }
    __syncthreads();
  }
  const unsigned long long t = sh.prefix;
  const long long want = sh.remaining;  // entries equal to t to keep, >= 1
  double ties = 0.0;
  for (int64_t e = tid; e < n; e += nt) ties += abs_key(v[e]) == t ? 1.0 : 0.0;
  ties = block_sum(ties, sh.red);
  if (tid == 0) {
    long long cut = 0;  // keep ties at flat index >= cut
    if ((long long)ties != want) {
      long long seen = 0;
      for (int64_t e = n - 1; e >= 0; --e)
        if (abs_key(v[e]) == t && ++seen == want) {
          cut = e;
          break;
        }
    }
    sh.cut = cut;
  }
  __syncthreads();
  const long long cut = sh.cut;
  for (int64_t e = tid; e < n; e += nt) {
    const unsigned long long key = abs_key(v[e]);
    x[e] = (key > t || (key == t && e >= cut)) ? v[e] : 0.0;
  }
}

// Per column: projection on {x >= 0, sum x = z} of u = v (or |v|, then
// times sign(v): soft_sparsity).  The threshold is the reference's c_K
// (tensorly/tenalg/proximal.py:592-631), reached by Michelot's fixed point
// tau = (sum of the entries above tau - z) / their count; for z <= 0 no k
// qualifies in the reference and the shift is the last c, (sum u - z) / rows.
__device__ void simplex_block(const ProxBufs& b, bool soft) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int64_t rows = b.rows;
  const int R = b.R;
  const double z = b.param;
  for (int c = wave; c < R; c += nw) {
    double s = 0.0;
    for (int64_t i = lane; i < rows; i += 64) {
      const double u = soft ? fabs(b.v[i * R + c]) : b.v[i * R + c];
      s += u;
    }
    s = wave_sum(s);
    double tau = (s - z) / (double)rows;
    if (z > 0.0) {
      double cnt_prev = (double)rows;
      for (int64_t it = 0; it <= rows; ++it) {
        double ss = 0.0, cnt = 0.0;
        for (int64_t i = lane; i < rows; i += 64) {
          const double u = soft ? fabs(b.v[i * R + c]) : b.v[i * R + c];
          if (u > tau) {
            ss += u;
            cnt += 1.0;
          }
        }
        ss = wave_sum(ss);
        cnt = wave_sum(cnt);
        if (cnt == cnt_prev || cnt == 0.0) break;
        tau = (ss - z) / cnt;
        cnt_prev = cnt;
      }
    }
    for (int64_t i = lane; i < rows; i += 64) {
      const double vv = b.v[i * R + c];
      const double u = soft ? fabs(vv) : vv;
      double o = fmax(u - tau, 0.0);
      if (soft) o *= (vv > 0.0) ? 1.0 : ((vv < 0.0) ? -1.0 : 0.0);
      b.x[i * R + c] = o;
    }
  }
}

// Per column: solve T x = v with T tridiagonal (2 lambda + 1 on the
// diagonal, -lambda beside it) by LU without pivoting (T is diagonally
// dominant, so partial pivoting never swaps), t1 holding the pivots
// (tensorly/tenalg/proximal.py:270-291).
__device__ void smoothness_block(const ProxBufs& b) {
  const int c = threadIdx.x;
  if (c >= b.R) return;
  const int R = b.R;
  const int64_t n = b.rows;
  const double lam = b.param, diag = 2.0 * lam + 1.0;
  double d = diag, y = b.v[c];
  b.t1[c] = d;
  b.x[c] = y;
  for (int64_t k = 1; k < n; ++k) {
    const double l = -lam / d;
    d = diag - l * (-lam);
    y = b.v[k * R + c] - l * y;
    b.t1[k * R + c] = d;
    b.x[k * R + c] = y;
  }
  double xn = b.x[(n - 1) * R + c] / d;
  b.x[(n - 1) * R + c] = xn;
  for (int64_t k = n - 2; k >= 0; --k) {
    xn = (b.x[k * R + c] - (-lam) * xn) / b.t1[k * R + c];
    b.x[k * R + c] = xn;
  }
}

// Per column (tensorly/tenalg/proximal.py:294-356), in the order j of the
// (flipped, when decreasing) column: y_j = max over i <= j of the mean of
// u_i..u_j, floored at 0 for j < rows - 1 (the reference's helper matrix
// holds zeros below its diagonal), then y_j = min(y_j, y_{j+1}) from the
// bottom up; out gets y at the unflipped positions.  cum and y: rows x R
// scratch.  Every thread of the block takes part.
__device__ void monotone_block(const double* v, double* out, double* cum, double* y, int64_t n, int R,
                               bool decreasing) {
  const int tid = threadIdx.x, nt = blockDim.x;
  auto row = [&](int64_t j) { return decreasing ? n - 1 - j : j; };
  if (tid < R) {
    double s = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      s += v[row(j) * R + tid];
      cum[j * R + tid] = s;
    }
  }
  __syncthreads();
  for (int64_t e = tid; e < n * R; e += nt) {
    const int64_t j = e / R;
    const int c = (int)(e % R);
    const double cj = cum[j * R + c];
    double best = j < n - 1 ? 0.0 : -INFINITY;
    best = fmax(best, cj / (double)(j + 1));
    for (int64_t i = 1; i <= j; ++i) best = fmax(best, (cj - cum[(i - 1) * R + c]) / (double)(j - i + 1));
    y[j * R + c] = best;
  }
  __syncthreads();
  if (tid < R) {
    double next = y[(n - 1) * R + tid];
    out[row(n - 1) * R + tid] = next;
    for (int64_t j = n - 2; j >= 0; --j) {
      double cur = y[j * R + tid];
      if (cur > next) cur = next;
      out[row(j) * R + tid] = cur;
      next = cur;
    }
  }
  __syncthreads();
}

// tensorly/tenalg/proximal.py:359-475: per column, the peak j among the
// entries at or above both monotone projections with the least
// sum_{k<j}|v_k - inc_k| + sum_{k>j}|v_k - dec_k|; non-peak entries count as
// the largest such cost over the whole matrix (first index on ties); the
// result is inc above the peak, v at it and dec below.
__device__ void unimodal_block(const ProxBufs& b, ProxShared& sh) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t n = b.rows;
  const int R = b.R;
  double* inc = b.t1;
  double* dec = b.t2;
  // x serves as the scratch of the two monotone passes, t3 then holds costs
  monotone_block(b.v, inc, b.t3, b.x, n, R, false);
  monotone_block(b.v, dec, b.t3, b.x, n, R, true);
  double colmax = 0.0;
  if (tid < R) {
    const int c = tid;
    double acc = 0.0;
    for (int64_t k = n - 1; k >= 0; --k) {
      const double a = fabs(b.v[k * R + c] - dec[k * R + c]);
      acc += a;
      b.t3[k * R + c] = acc - a;  // sum over the entries below k
    }
    acc = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      const double vj = b.v[j * R + c];
      const double a = fabs(vj - inc[j * R + c]);
      acc += a;
      const bool peak = (vj - dec[j * R + c] >= 0.0) && (vj - inc[j * R + c] >= 0.0);
      const double cost = (acc - a) + b.t3[j * R + c];
      b.t3[j * R + c] = peak ? cost : -1.0;  // costs are >= 0: -1 marks a non-peak
      if (peak) colmax = fmax(colmax, cost);
    }
  }
  const double M = block_max(colmax, sh.red);  // >= 0: non-peaks contribute 0
  if (tid < R) {
    const int c = tid;
    int64_t arg = 0;
    double best = INFINITY;
    for (int64_t j = 0; j < n; ++j) {
      double cost = b.t3[j * R + c];
      if (cost < 0.0) cost = M;
      if (cost < best) {
        best = cost;
        arg = j;
      }
    }
    sh.jstar[c] = (int)arg;
  }
  __syncthreads();
  for (int64_t e = tid; e < n * R; e += nt) {
    const int64_t j = e / R;
    const int c = (int)(e % R);
    const int64_t js = sh.jstar[c];
    b.x[e] = j < js ? inc[e] : (j == js ? b.v[e] : dec[e]);
  }
}

// x = prox(v) for constraint b.kind (tensorly/tenalg/proximal.py:153-267).
// Every thread of the block calls it; ends with a barrier.
__device__ void prox_block(const ProxBufs& b, ProxShared& sh) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t n = b.rows * b.R;
  const double p = b.param;
  switch (b.kind) {
    case TLA_CON_NON_NEGATIVE: {
      // np.clip(v, 0, max(v)): when every entry is negative the bounds cross
      // and numpy returns the upper bound everywhere
      double m = -INFINITY;
      for (int64_t e = tid; e < n; e += nt) m = fmax(m, b.v[e]);
      m = block_max(m, sh.red);
      for (int64_t e = tid; e < n; e += nt) b.x[e] = fmin(fmax(b.v[e], 0.0), m);
      break;
    }
    case TLA_CON_L1:
      for (int64_t e = tid; e < n; e += nt) {
        const double v = b.v[e];
        const double a = fmax(fabs(v) - p, 0.0);
        b.x[e] = ((v > 0.0) ? 1.0 : ((v < 0.0) ? -1.0 : 0.0)) * a;
      }
      break;
    case TLA_CON_L2: {
      double s = 0.0;
      for (int64_t e = tid; e < n; e += nt) s = fma(b.v[e], b.v[e], s);
      const double nrm = sqrt(block_sum(s, sh.red));
      const double big = nrm > p ? nrm : p;
      for (int64_t e = tid; e < n; e += nt) b.x[e] = b.v[e] - b.v[e] * p / big;
      break;
    }
    case TLA_CON_L2_SQUARE:
      for (int64_t e = tid; e < n; e += nt) b.x[e] = b.v[e] / (1.0 + 2.0 * p);
      break;
    case TLA_CON_NORMALIZE: {
      double m = 0.0;
      for (int64_t e = tid; e < n; e += nt) m = fmax(m, fabs(b.v[e]));
      m = block_max(m, sh.red);
      for (int64_t e = tid; e < n; e += nt) b.x[e] = b.v[e] / m;
      break;
    }
    case TLA_CON_HARD_SPARSITY:
      hard_threshold_block(b.v, b.x, n, p, sh);
      break;
    case TLA_CON_NORMALIZED_SPARSITY: {
      hard_threshold_block(b.v, b.x, n, p, sh);
      __syncthreads();
      double s = 0.0;
      for (int64_t e = tid; e < n; e += nt) s = fma(b.x[e], b.x[e], s);
      const double nrm = sqrt(block_sum(s, sh.red));
      for (int64_t e = tid; e < n; e += nt) b.x[e] = b.x[e] / nrm;
      break;
    }
    case TLA_CON_SIMPLEX:
      simplex_block(b, false);
      break;
    case TLA_CON_SOFT_SPARSITY:
      simplex_block(b, true);
      break;
    case TLA_CON_SMOOTHNESS:
      smoothness_block(b);
      break;
    case TLA_CON_MONOTONICITY:
      monotone_block(b.v, b.x, b.t3, b.t2, b.rows, b.R, false);
      break;
    case TLA_CON_UNIMODALITY:
      unimodal_block(b, sh);
      break;
    default:  // TLA_CON_NONE
      for (int64_t e = tid; e < n; e += nt) b.x[e] = b.v[e];
      break;
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ double ldd(const void* p, int64_t i) {
  return (double)((const T*)p)[i];
}

// ---------------------------------------------------------------------------
struct ProxArgs {
  const void* in;
  void* out;
  double* ws;  // 5 rows x R FP64 matrices
  int64_t rows;
  int R;
  int kind;
  double param;
};

template <typename T>
__global__ __launch_bounds__(kAdmmThreads) void prox_kernel(ProxArgs a) {
  __shared__ ProxShared sh;
  const int64_t n = a.rows * a.R;
  double* V = a.ws;
  ProxBufs b{V, V + n, V + 2 * n, V + 3 * n, V + 4 * n, a.rows, a.R, a.kind, a.param};
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) V[e] = ldd<T>(a.in, e);
  __syncthreads();
  prox_block(b, sh);
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) ((T*)a.out)[e] = (T)b.x[e];
}

// ---------------------------------------------------------------------------
struct AdmmArgs {
  const double* grams[kMaxOrder];
  int ndim, mode;
  const double* utu;  // R x R, or NULL: Hadamard product of the other Grams
  int R;
  int64_t rows;
  const void* M;   // UtM, rows x R (dtype)
  void* factor;    // x, rows x R (dtype), in / out
  double* xsplit;  // x_split^T, rows x R FP64 out
  double* dual;    // rows x R FP64 in / out
  int n_iter_max;
  double tol;
  int kind;
  double param;
  double* ws;  // FP64: X, V, XO, T1, T2, T3 (rows x R each), then 2 R x R
  int* status;
  int* iters_out;
  double* norms_out;  // [||x - x_split^T||, ||x||] of the last iteration
  double* iprod_out;  // sum(UtM * x) with x as stored
};

template <typename T, int RT>
__global__ __launch_bounds__(kAdmmThreads) void admm_kernel(AdmmArgs a) {
  __shared__ double Minv[RT * RT];
  __shared__ ProxShared sh;
  __shared__ int piv, flag, stop;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int R = a.R;
  const int64_t rows = a.rows, n = rows * R;
  double* X = a.ws;
  double* V = X + n;
  double* XO = V + n;
  double* T1 = XO + n;
  double* T2 = T1 + n;
  double* T3 = T2 + n;
  double* A = T3 + n;      // UtU + rho I, destroyed by the elimination
  double* Ai = A + R * R;  // its inverse
  double* XS = a.xsplit;
  double* D = a.dual;
  // UtU and rho = trace(UtU) / R
  for (int e = tid; e < R * R; e += nt) {
    double u;
    if (a.utu) {
      u = a.utu[e];
    } else {
      u = 1.0;
      for (int m = 0; m < a.ndim; ++m)
        if (m != a.mode) u *= a.grams[m][e];
    }
    A[e] = u;
  }
  __syncthreads();
  double tr = 0.0;
  for (int r = 0; r < R; ++r) tr += A[r * R + r];
  const double rho = tr / (double)R;
  __syncthreads();
  for (int r = tid; r < R; r += nt) A[r * R + r] += rho;
  __syncthreads();
  const bool ok = lds_gj_inverse(A, Ai, R, &piv, &flag);
  if (!ok) {
    if (tid == 0) {
      if (a.status) *a.status = 2;
      if (a.iters_out) *a.iters_out = 0;
    }
    return;
  }
  for (int e = tid; e < RT * RT; e += nt) {
    const int r = e / RT, c = e % RT;
    Minv[e] = (r < R && c < R) ? Ai[r * R + c] : 0.0;
  }
  for (int64_t e = tid; e < n; e += nt) X[e] = ldd<T>(a.factor, e);
  __syncthreads();
  ProxBufs pb{V, X, T1, T2, T3, rows, R, a.kind, a.param};
  double nrm_dual = 0.0, nrm_x = 0.0;
  int it = 0;
  for (; it < a.n_iter_max; ++it) {
    // x_split^T = (UtM + rho (x + dual)) (UtU + rho I)^-1, one row per thread
    for (int64_t i = tid; i < rows; i += nt) {
      const double* Mi = Minv;
      asm volatile("" : "+v"(Mi));
      double t[RT];
#pragma unroll
      for (int k = 0; k < RT; ++k)
        t[k] = k < R ? ldd<T>(a.M, i * R + k) + rho * (X[i * R + k] + D[i * R + k]) : 0.0;
      for (int c = 0; c < R; ++c) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < RT; ++k) s = fma(t[k], Mi[k * RT + c], s);
        XS[i * R + c] = s;
        V[i * R + c] = s - D[i * R + c];
        XO[i * R + c] = X[i * R + c];
      }
    }
    __syncthreads();
    prox_block(pb, sh);
    double s_dual = 0.0, s_x = 0.0, s_pr = 0.0, s_d = 0.0;
    for (int64_t e = tid; e < n; e += nt) {
      const double x = X[e], xs = XS[e];
      const double d = D[e] + x - xs;
      D[e] = d;
      const double r = x - xs, pr = x - XO[e];
      s_dual = fma(r, r, s_dual);
      s_x = fma(x, x, s_x);
      s_pr = fma(pr, pr, s_pr);
      s_d = fma(d, d, s_d);
    }
    s_dual = block_sum(s_dual, sh.red);
    s_x = block_sum(s_x, sh.red);
    s_pr = block_sum(s_pr, sh.red);
    s_d = block_sum(s_d, sh.red);
    nrm_dual = sqrt(s_dual);
    nrm_x = sqrt(s_x);
    if (tid == 0) stop = (nrm_dual < a.tol * nrm_x && sqrt(s_pr) < a.tol * sqrt(s_d)) ? 1 : 0;
    __syncthreads();
    const bool done = stop != 0;
    __syncthreads();
    if (done) {
      ++it;
      break;
    }
  }
  double ip = 0.0;
  for (int64_t e = tid; e < n; e += nt) {
    const T o = (T)X[e];
    ((T*)a.factor)[e] = o;
    ip = fma(ldd<T>(a.M, e), (double)o, ip);
  }
  ip = block_sum(ip, sh.red);
  if (tid == 0) {
    if (a.status) *a.status = 0;
    if (a.iters_out) *a.iters_out = it;
    if (a.norms_out) {
      a.norms_out[0] = nrm_dual;
      a.norms_out[1] = nrm_x;
    }
    if (a.iprod_out) *a.iprod_out = ip;
  }
}

}  // namespace tla

using namespace tla;

static bool valid_kind(int k) { return k >= TLA_CON_NONE && k <= TLA_CON_HARD_SPARSITY; }

extern "C" size_t tla_prox_workspace_size(int64_t rows, int cols) {
  if (rows < 1 || cols < 1) return 0;
  return size_t(5) * rows * cols * sizeof(double);
}

extern "C" int tla_prox(int dtype, const void* in, void* out, int64_t rows, int cols, int constraint, double param,
                        void* workspace, size_t workspace_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "prox: bad dtype %d", dtype);
  TLA_REQUIRE(rows >= 1 && cols >= 1 && cols <= kMaxRank, "prox: shape (%lld, %d) needs 1 <= cols <= %d",
              (long long)rows, cols, kMaxRank);
  TLA_REQUIRE(valid_kind(constraint), "prox: unknown constraint %d", constraint);
  TLA_REQUIRE(in && out && workspace, "prox: NULL pointer");
  TLA_REQUIRE(workspace_bytes >= tla_prox_workspace_size(rows, cols), "prox: workspace too small");
  ProxArgs a{in, out, (double*)workspace, rows, cols, constraint, param};
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) prox_kernel<float><<<1, kAdmmThreads, 0, st>>>(a);
  else prox_kernel<double><<<1, kAdmmThreads, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" size_t tla_admm_workspace_size(int64_t rows, int rank) {
  if (rows < 1 || rank < 1) return 0;
  return (size_t(6) * rows * rank + size_t(2) * rank * rank) * sizeof(double);
}

extern "C" int tla_admm_update(int dtype, void* factor, double* x_split, double* dual, const void* mttkrp,
                               int64_t rows, int rank, const double* utu, const double* const* grams, int ndim,
                               int mode, int n_iter_max, double tol, int constraint, double param, int* status_out,
                               int* iters_out, double* norms_out, double* iprod_out, void* workspace,
                               size_t workspace_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "admm: bad dtype %d", dtype);
  TLA_REQUIRE(rows >= 1 && rank >= 1 && rank <= kMaxRank, "admm: shape (%lld, %d) needs 1 <= rank <= %d",
              (long long)rows, rank, kMaxRank);
  TLA_REQUIRE(n_iter_max >= 1, "admm: n_iter_max must be >= 1");
  TLA_REQUIRE(valid_kind(constraint), "admm: unknown constraint %d", constraint);
  TLA_REQUIRE(factor && x_split && dual && mttkrp && workspace, "admm: NULL pointer");
  TLA_REQUIRE(workspace_bytes >= tla_admm_workspace_size(rows, rank), "admm: workspace too small");
  AdmmArgs a{};
  if (!utu) {
    TLA_REQUIRE(grams && ndim >= 2 && ndim <= kMaxOrder && mode >= 0 && mode < ndim,
                "admm: UtU from the Grams needs ndim and mode");
    for (int m = 0; m < ndim; ++m) {
      TLA_REQUIRE(m == mode || grams[m], "admm: gram %d is NULL", m);
      a.grams[m] = grams[m];
    }
  }
  a.ndim = ndim; a.mode = mode; a.utu = utu; a.R = rank; a.rows = rows;
  a.M = mttkrp; a.factor = factor; a.xsplit = x_split; a.dual = dual;
  a.n_iter_max = n_iter_max; a.tol = tol; a.kind = constraint; a.param = param;
  a.ws = (double*)workspace; a.status = status_out; a.iters_out = iters_out; a.norms_out = norms_out;
  a.iprod_out = iprod_out;
  hipStream_t st = (hipStream_t)stream;
  const int RT = rank <= 16 ? 16 : (rank <= 32 ? 32 : (rank <= 48 ? 48 : 64));
#define TLA_ADMM_CASE(RTV)                                                          \
  case RTV:                                                                         \
    if (dtype == TLA_F32) admm_kernel<float, RTV><<<1, kAdmmThreads, 0, st>>>(a);   \
    else admm_kernel<double, RTV><<<1, kAdmmThreads, 0, st>>>(a);                   \
    break;
  switch (RT) {
    TLA_ADMM_CASE(16)
    TLA_ADMM_CASE(32)
    TLA_ADMM_CASE(48)
    TLA_ADMM_CASE(64)
  }
#undef TLA_ADMM_CASE
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
