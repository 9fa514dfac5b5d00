// FP64 symmetric eigensolver for the leading eigenpairs of a Gram matrix,
// and the mode-n Gram G = unfold(X, n) unfold(X, n)^T.
//
// These replace the LAPACK SVDs of tensorly/tenalg/svd.py:211-235
// (truncated_svd) on the Tucker / CP-init path: the left singular vectors of
// A are the eigenvectors of A A^T and the singular values the square roots
// of its eigenvalues.
//
//  * n <= 64: cyclic Jacobi on the whole matrix in LDS (round-robin pair
//    ordering, all n/2 rotations of a round applied in parallel).
//  * n  > 64: block subspace iteration with Rayleigh-Ritz.  Block b = 64
//    for k <= 56 (Ritz matrix and rotations in LDS); otherwise b =
//    min(n, 128) for k <= 120 or n <= 128 (Ritz matrix in LDS, rotations in
//    the workspace; b = n makes the first Rayleigh-Ritz step exact).  Each
//    step: Z = G V (FP64), H = V^T Z, Jacobi(H) = W diag(t) W^T (warm: the
//    previous Ritz basis nearly diagonalises H), residuals of the k leading
//    Ritz pairs, V <- orth(Z W) by shifted Cholesky-QR (three passes).
//    Converged when every leading residual is below tol * t_1.
// One workgroup per problem; the n x b blocks live in the caller's
// workspace (L2-resident).
#include <math.h>

#include <algorithm>

#include "common.h"

namespace tla {

constexpr int kEigThreads = 512;
constexpr int kJac = 64;          // max Jacobi size (LDS)
constexpr int kEigCapLarge = 128;  // larger subspace block

// ---------------------------------------------------------------------------
// Jacobi on an m x m symmetric matrix A (row-major, leading dimension LD),
// accumulating rotations into W (same layout, initialised by the caller).
// A lives in LDS; W may be in LDS or in global memory (the LD = 128 case,
// where A alone takes 128 KB of LDS).  Returns the number of sweeps.
// ---------------------------------------------------------------------------
template <int LD>
__device__ int lds_jacobi(double* A, double* W, int m, int max_sweeps, double* cs, int* pr, double* scratch) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int me = m + (m & 1);  // even number of slots (one dummy)
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    // off-diagonal and total Frobenius norms
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < m * m; e += nt) {
      const int i = e / m, j = e % m;
      const double v = A[i * LD + j];
      tot += v * v;
      if (i != j) off += v * v;
    }
    off = wave_sum(off);
    tot = wave_sum(tot);
    if ((tid & 63) == 0) { scratch[tid >> 6] = off; scratch[16 + (tid >> 6)] = tot; }
    __syncthreads();
    if (tid == 0) {
      double o = 0.0, t = 0.0;
      for (int w = 0; w < nt / 64; ++w) { o += scratch[w]; t += scratch[16 + w]; }
      scratch[32] = o;
      scratch[33] = t;
    }
    __syncthreads();
    if (!(scratch[32] > 1e-30 * scratch[33]) || scratch[33] == 0.0) break;
    for (int round = 0; round < me - 1; ++round) {
      // round-robin pairing: slot 0 fixed, others rotate
      if (tid < me / 2) {
        const int a = tid;
        int p = (a == 0) ? 0 : 1 + (a - 1 + round) % (me - 1);
        int q = 1 + (me - 1 - a - 1 + round) % (me - 1);
        if (a == 0) q = 1 + (me - 2 + round) % (me - 1);
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, s = 0.0;
        if (q < m) {
          const double apq = A[p * LD + q];
          if (fabs(apq) > 1e-300) {
            const double app = A[p * LD + p], aqq = A[q * LD + q];
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        pr[2 * a] = p;
        pr[2 * a + 1] = q;
        cs[2 * a] = c;
        cs[2 * a + 1] = s;
      }
      __syncthreads();
      // rows: A[p,:], A[q,:]
      const int npairs = me / 2;
      for (int e = tid; e < npairs * m; e += nt) {
        const int a = e / m, j = e % m;
        const int p = pr[2 * a], q = pr[2 * a + 1];
        if (q < m) {
          const double c = cs[2 * a], s = cs[2 * a + 1];
          const double x = A[p * LD + j], y = A[q * LD + j];
          A[p * LD + j] = c * x - s * y;
          A[q * LD + j] = s * x + c * y;
        }
      }
      __syncthreads();
      // columns of A and W
      for (int e = tid; e < npairs * m; e += nt) {
        const int a = e / m, i = e % m;
        const int p = pr[2 * a], q = pr[2 * a + 1];
        if (q < m) {
          const double c = cs[2 * a], s = cs[2 * a + 1];
          double x = A[i * LD + p], y = A[i * LD + q];
          A[i * LD + p] = c * x - s * y;
          A[i * LD + q] = s * x + c * y;
          x = W[i * LD + p];
          y = W[i * LD + q];
          W[i * LD + p] = c * x - s * y;
          W[i * LD + q] = s * x + c * y;
        }
      }
      __syncthreads();
    }
  }
  return sweep;
}

// order[r] = index of the r-th largest diagonal entry of A (ties by index)
template <int LD>
__device__ void sort_desc(const double* A, int m, int* order) {
  const int tid = threadIdx.x;
  if (tid < m) {
    const double v = A[tid * LD + tid];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double u = A[j * LD + j];
      if (u > v || (u == v && j < tid)) ++rank;
    }
    order[rank] = tid;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
struct EigArgs {
  const double* G;
  int n, k;
  const double* Q0;  // warm start (row-major, ld q0_ld) or nullptr; its first k0 columns are used
  int k0, q0_ld;
  double* evals;     // k
  double* evecs;     // n x k row-major
  double tol;
  int max_iter;
  int* iters;
  int b;      // subspace block: columns in use, <= the kernel's capacity
  double* V;  // n x capacity
  double* Z;  // n x capacity
  double* Y;  // n x capacity
  double* U;  // n x capacity
  double* W;  // capacity x capacity (Ritz rotations; used when capacity = 128)
};

__device__ __forceinline__ double hash_unit(uint32_t i, uint32_t j) {
  uint32_t h = i * 0x9E3779B1u ^ (j + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
  return (double)h * (1.0 / 4294967296.0) - 0.5;
}

// Shifted Cholesky-QR of the n x b block M (global, row-major ld LD) in
// place; S (LDS, ld LD) scratch.
template <int LD>
__device__ void block_orth(double* M, int n, int b, double* S, double* scratch, int* flag) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int pass = 0; pass < 3; ++pass) {
    // S = M^T M
    for (int e = tid; e < b * b; e += nt) {
      const int i = e / b, j = e % b;
      if (j < i) continue;
      double s = 0.0;
      for (int r = 0; r < n; ++r) s = fma(M[r * LD + i], M[r * LD + j], s);
      S[i * LD + j] = s;
      S[j * LD + i] = s;
    }
    __syncthreads();
    if (tid == 0) {
      double tr = 0.0;
      for (int i = 0; i < b; ++i) tr = fmax(tr, S[i * LD + i]);
      scratch[40] = tr;
    }
    __syncthreads();
    const double shift = (pass == 0) ? 1e-13 * scratch[40] * (double)n : 0.0;
    // Cholesky S + shift I = R^T R (R upper, stored in S upper)
    if (tid == 0) *flag = 0;
    __syncthreads();
    for (int j = 0; j < b; ++j) {
      if (tid == 0) {
        double d = S[j * LD + j] + shift;
        for (int l = 0; l < j; ++l) d -= S[l * LD + j] * S[l * LD + j];
        if (!(d > 0.0)) { d = 1e-300; *flag = 1; }
        S[j * LD + j] = sqrt(d);
      }
      __syncthreads();
      for (int i = j + 1 + tid; i < b; i += nt) {
        double v = S[j * LD + i];
        for (int l = 0; l < j; ++l) v -= S[l * LD + j] * S[l * LD + i];
        S[j * LD + i] = v / S[j * LD + j];
      }
      __syncthreads();
    }
    // M = M R^-1 (row-wise forward substitution)
    for (int r = tid; r < n; r += nt) {
      double* row = M + r * LD;
      for (int j = 0; j < b; ++j) {
        double v = row[j];
        for (int l = 0; l < j; ++l) v -= row[l] * S[l * LD + j];
        row[j] = v / S[j * LD + j];
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kEigThreads) void sym_eig_small_kernel(EigArgs a) {
  __shared__ double A[kJac * kJac];
  __shared__ double W[kJac * kJac];
  __shared__ double cs[kJac + 2];
  __shared__ int pr[kJac + 2];
  __shared__ double scratch[64];
  __shared__ int order[kJac];
  const int n = a.n, tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    A[i * kJac + j] = 0.5 * (a.G[i * n + j] + a.G[j * n + i]);
    W[i * kJac + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int sweeps = lds_jacobi<kJac>(A, W, n, 60, cs, pr, scratch);
  sort_desc<kJac>(A, n, order);
  for (int e = tid; e < n * a.k; e += blockDim.x) {
    const int i = e / a.k, r = e % a.k;
    a.evecs[i * a.k + r] = W[i * kJac + order[r]];
  }
  if (tid < a.k) a.evals[tid] = A[order[tid] * kJac + order[tid]];
  if (tid == 0 && a.iters) *a.iters = sweeps;
}

// Block subspace iteration with Rayleigh-Ritz.  CAP is the block capacity:
// 64 keeps H and the rotations W in LDS; 128 keeps H (128 KB) in LDS and W
// in the workspace.  a.b <= CAP columns are used; with a.b == n the first
// Rayleigh-Ritz step is already exact.
template <int CAP>
__global__ __launch_bounds__(kEigThreads) void sym_eig_subspace_kernel(EigArgs a) {
  constexpr int LD = CAP;
  __shared__ double H[LD * LD];
  __shared__ double Wl[CAP == 64 ? LD * LD : 1];
  __shared__ double cs[LD + 2];
  __shared__ int pr[LD + 2];
  __shared__ double scratch[64];
  __shared__ double resid[LD];
  __shared__ int order[LD];
  __shared__ double theta[LD];
  __shared__ int flag;
  const int n = a.n, k = a.k, b = a.b, tid = threadIdx.x, nt = blockDim.x;
  double* W = (CAP == 64) ? Wl : a.W;
  double* V = a.V;
  double* Z = a.Z;
  double* Y = a.Y;
  double* U = a.U;
  // V0: warm start columns, then deterministic pseudo-random columns
  for (int e = tid; e < n * b; e += nt) {
    const int i = e / b, j = e % b;
    V[i * LD + j] = (a.Q0 && j < a.k0) ? a.Q0[(int64_t)i * a.q0_ld + j] : hash_unit(i, j);
  }
  __syncthreads();
  block_orth<LD>(V, n, b, H, scratch, &flag);
  int it = 0;
  for (; it < a.max_iter; ++it) {
    // Z = G V
    for (int e = tid; e < n * b; e += nt) {
      const int i = e / b, j = e % b;
      const double* g = a.G + (int64_t)i * n;
      double s = 0.0;
      for (int l = 0; l < n; ++l) s = fma(g[l], V[l * LD + j], s);
      Z[i * LD + j] = s;
    }
    __syncthreads();
    // H = V^T Z (symmetrised), W = I
    for (int e = tid; e < b * b; e += nt) {
      const int i = e / b, j = e % b;
      if (j < i) continue;
      double s = 0.0, t = 0.0;
      for (int r = 0; r < n; ++r) {
        s = fma(V[r * LD + i], Z[r * LD + j], s);
        t = fma(V[r * LD + j], Z[r * LD + i], t);
      }
      H[i * LD + j] = H[j * LD + i] = 0.5 * (s + t);
    }
    for (int e = tid; e < b * b; e += nt) W[(e / b) * LD + e % b] = ((e / b) == (e % b)) ? 1.0 : 0.0;
    __syncthreads();
    lds_jacobi<LD>(H, W, b, 30, cs, pr, scratch);
    sort_desc<LD>(H, b, order);
    if (tid < b) theta[tid] = H[order[tid] * LD + order[tid]];
    __syncthreads();
    // Y = V W (Ritz vectors, sorted by Ritz value), U = Z W = G Y
    for (int e = tid; e < n * b; e += nt) {
      const int i = e / b, j = e % b;
      const int oj = order[j];
      double y = 0.0, z = 0.0;
      for (int l = 0; l < b; ++l) {
        const double w = W[l * LD + oj];
        y = fma(V[i * LD + l], w, y);
        z = fma(Z[i * LD + l], w, z);
      }
      Y[i * LD + j] = y;
      U[i * LD + j] = z;
    }
    __syncthreads();
    // residual norms of the leading k pairs
    for (int r = tid; r < k; r += nt) {
      double r2 = 0.0;
      for (int i = 0; i < n; ++i) {
        const double d = U[i * LD + r] - theta[r] * Y[i * LD + r];
        r2 += d * d;
      }
      resid[r] = sqrt(r2);
    }
    __syncthreads();
    if (tid == 0) {
      double worst = 0.0;
      for (int r = 0; r < k; ++r) worst = fmax(worst, resid[r]);
      scratch[63] = (worst <= a.tol * fmax(fabs(theta[0]), 1e-300)) ? 1.0 : 0.0;
    }
    __syncthreads();
    const bool done = scratch[63] != 0.0;
    __syncthreads();
    if (done || it + 1 == a.max_iter) break;
    // V <- orth(Z W)
    for (int e = tid; e < n * b; e += nt) V[(e / b) * LD + e % b] = U[(e / b) * LD + e % b];
    __syncthreads();
    block_orth<LD>(V, n, b, H, scratch, &flag);
  }
  for (int e = tid; e < n * k; e += nt) {
    const int i = e / k, r = e % k;
    a.evecs[i * k + r] = Y[i * LD + r];
  }
  for (int r = tid; r < k; r += nt) a.evals[r] = theta[r];
  if (tid == 0 && a.iters) *a.iters = it + 1;
}

// ---------------------------------------------------------------------------
// Mode Gram: G[i, j] = sum_{a,b} X[a, i, b] X[a, j, b], FP64 accumulation.
// Block (ti, tj) computes a 32 x 32 tile for one chunk of the (a, b) pairs;
// partial tiles are summed in chunk order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void mode_gram_kernel(const T* __restrict__ X, int64_t Pa, int64_t In, int64_t Pb,
                                                        int64_t chunk, double* __restrict__ partial) {
  __shared__ double Ti[32][65];
  __shared__ double Tj[32][65];
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (tj < ti) return;
  const int tid = threadIdx.x;
  const int64_t q0 = (int64_t)blockIdx.z * chunk;
  const int64_t Q = Pa * Pb;
  const int64_t q1 = min(Q, q0 + chunk);
  // each thread owns 4 outputs (ri, cj..cj+3)
  const int ri = tid / 8, cj = (tid % 8) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t qb = q0; qb < q1; qb += 64) {
    for (int e = tid; e < 32 * 64; e += 256) {
      const int r = e / 64, qq = e % 64;
      const int64_t q = qb + qq;
      const int64_t gi = (int64_t)ti * 32 + r, gj = (int64_t)tj * 32 + r;
      double vi = 0.0, vj = 0.0;
      if (q < q1) {
        const int64_t aa = q / Pb, bb = q - aa * Pb;
        if (gi < In) vi = (double)X[(aa * In + gi) * Pb + bb];
        if (gj < In) vj = (double)X[(aa * In + gj) * Pb + bb];
      }
      Ti[r][qq] = vi;
      Tj[r][qq] = vj;
    }
    __syncthreads();
    for (int qq = 0; qq < 64; ++qq) {
      const double x = Ti[ri][qq];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] = fma(x, Tj[cj + c][qq], acc[c]);
    }
    __syncthreads();
  }
  const int64_t nt = (int64_t)gridDim.x;
  double* out = partial + (int64_t)blockIdx.z * In * In;
  (void)nt;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int64_t i = (int64_t)ti * 32 + ri, j = (int64_t)tj * 32 + cj + c;
    if (i < In && j < In) {
      out[i * In + j] = acc[c];
      out[j * In + i] = acc[c];
    }
  }
}

__global__ void sum_gram_partials_kernel(const double* __restrict__ partial, int64_t nn, int splits,
                                         double* __restrict__ G) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  double s = 0.0;
  for (int k = 0; k < splits; ++k) s += partial[(int64_t)k * nn + e];
  G[e] = s;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static int gram_splits(int64_t In, int64_t Q) {
  const int64_t tiles = ((In + 31) / 32) * ((In + 31) / 32 + 1) / 2;
  int64_t s = std::max<int64_t>(1, 1024 / std::max<int64_t>(tiles, 1));
  s = std::min<int64_t>(s, (Q + 63) / 64);
  s = std::min<int64_t>(s, 64);
  return (int)std::max<int64_t>(1, s);
}

}  // namespace tla

using namespace tla;

extern "C" size_t tla_mode_gram64_workspace_size(int ndim, const int64_t* shape, int mode) {
  if (mode < 0 || mode >= ndim) return 0;
  int64_t In = shape[mode], Q = 1;
  for (int m = 0; m < ndim; ++m)
    if (m != mode) Q *= shape[m];
  return align256(size_t(gram_splits(In, Q)) * In * In * 8);
}

extern "C" int tla_mode_gram64(int dtype, const void* X, int ndim, const int64_t* shape, int mode, double* G,
                               void* ws, size_t ws_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "mode_gram64: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 1 && ndim <= kMaxOrder && mode >= 0 && mode < ndim, "mode_gram64: bad mode");
  TLA_REQUIRE(ws_bytes >= tla_mode_gram64_workspace_size(ndim, shape, mode), "mode_gram64: workspace too small");
  int64_t Pa = 1, Pb = 1;
  for (int m = 0; m < mode; ++m) Pa *= shape[m];
  for (int m = mode + 1; m < ndim; ++m) Pb *= shape[m];
  const int64_t In = shape[mode];
  const int splits = gram_splits(In, Pa * Pb);
  const int64_t Q = Pa * Pb;
  const int64_t chunk = ((Q + splits - 1) / splits + 63) / 64 * 64;
  const int nti = ceil_div(In, 32);
  hipStream_t st = (hipStream_t)stream;
  double* partial = (double*)ws;
  dim3 grid(nti, nti, splits);
  if (dtype == TLA_F32)
    mode_gram_kernel<float><<<grid, 256, 0, st>>>((const float*)X, Pa, In, Pb, chunk, partial);
  else
    mode_gram_kernel<double><<<grid, 256, 0, st>>>((const double*)X, Pa, In, Pb, chunk, partial);
  TLA_LAUNCH_CHECK();
  const int64_t nn = In * In;
  sum_gram_partials_kernel<<<ceil_div(nn, 256), 256, 0, st>>>(partial, nn, splits, G);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" size_t tla_sym_eig_workspace_size(int n) {
  if (n <= kJac) return 0;
  // four n x 128 blocks and the 128 x 128 rotations (sized for either block)
  return 4 * align256(size_t(n) * kEigCapLarge * 8) + align256(size_t(kEigCapLarge) * kEigCapLarge * 8);
}

extern "C" int tla_sym_eig(const double* G, int n, int k, const double* Q0, int k0, double* evals, double* evecs,
                           double tol, int max_iter, int* iters_out, void* ws, size_t ws_bytes, void* stream) {
  TLA_REQUIRE(n >= 1, "sym_eig: n = %d", n);
  TLA_REQUIRE(k >= 1 && k <= n, "sym_eig: k = %d not in [1, n = %d]", k, n);
  TLA_REQUIRE(G && evals && evecs, "sym_eig: NULL pointer");
  TLA_REQUIRE(!Q0 || k0 >= 1, "sym_eig: warm start with k0 = %d columns", k0);
  TLA_REQUIRE(ws_bytes >= tla_sym_eig_workspace_size(n), "sym_eig: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  EigArgs a{};
  a.G = G; a.n = n; a.k = k; a.Q0 = Q0; a.q0_ld = k0;
  a.evals = evals; a.evecs = evecs; a.tol = tol > 0 ? tol : 1e-11; a.max_iter = max_iter > 0 ? max_iter : 200;
  a.iters = iters_out;
  if (n <= kJac) {
    sym_eig_small_kernel<<<1, kEigThreads, 0, st>>>(a);
    TLA_LAUNCH_CHECK();
    return TLA_OK;
  }
  const bool small_block = k <= 56;
  if (!small_block && !(k <= kEigCapLarge - 8 || n <= kEigCapLarge)) {
    set_error("sym_eig: k = %d > %d leading eigenpairs of an n = %d > %d matrix is not supported", k,
              kEigCapLarge - 8, n, kEigCapLarge);
    return TLA_ERR_UNSUPPORTED;
  }
  const int cap = small_block ? 64 : kEigCapLarge;
  const size_t blk = align256(size_t(n) * cap * 8);
  a.b = small_block ? 64 : std::min(n, kEigCapLarge);
  a.k0 = Q0 ? std::min(k0, a.b) : 0;
  a.V = (double*)ws;
  a.Z = (double*)((char*)ws + blk);
  a.Y = (double*)((char*)ws + 2 * blk);
  a.U = (double*)((char*)ws + 3 * blk);
  a.W = (double*)((char*)ws + 4 * blk);
  if (small_block)
    sym_eig_subspace_kernel<64><<<1, kEigThreads, 0, st>>>(a);
  else
    sym_eig_subspace_kernel<kEigCapLarge><<<1, kEigThreads, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
