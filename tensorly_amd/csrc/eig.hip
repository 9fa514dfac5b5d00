// FP64 symmetric eigensolver for the leading eigenpairs of a Gram matrix,
// and the mode-n Gram G = unfold(X, n) unfold(X, n)^T.
//
// These replace the LAPACK SVDs of tensorly/tenalg/svd.py:211-235
// (truncated_svd) on the Tucker / CP-init path: the left singular vectors of
// A are the eigenvectors of A A^T and the singular values the square roots
// of its eigenvalues.
//
//  * n <= 128: cyclic Jacobi on the whole matrix (round-robin pair
//    ordering, all n/2 rotations of a round applied in parallel); the matrix
//    in LDS with an odd leading dimension, the rotations in LDS (n <= 64) or
//    in the workspace.
//  * n  > 128: block subspace iteration with Rayleigh-Ritz, block b = 48 for
//    k <= 40, 64 for k <= 56 or 128 for k <= 120.  Each step: Z = G V, H = V^T Z on FP64
//    MFMA, Jacobi(H) = W diag(t) W^T (warm: the previous Ritz basis nearly
//    diagonalises H), Ritz vectors Y = V W and U = Z W = G Y, residuals of
//    the k leading pairs, V <- orth(U) by two shifted Cholesky-QR passes
//    (S = M^T M on MFMA, forward elimination of [S | I] in registers).
//    Converged when every leading residual is below tol * t_1.
// One workgroup per problem; the padded np x b blocks live in the caller's
// workspace (L2-resident).
#include <math.h>

#include <algorithm>

#include "common.h"

namespace tla {

constexpr int kEigThreads = 512;
constexpr int kEigWaves = kEigThreads / 64;
constexpr int kJac = 64;           // n <= kJac: Jacobi with A and W in LDS
constexpr int kJacLarge = 128;     // n <= kJacLarge: Jacobi with A in LDS, W in the workspace
constexpr int kEigCapLarge = 128;  // larger subspace block

typedef double doublex4 __attribute__((ext_vector_type(4)));

// D = A B + C on FP64 MFMA (16 x 4 by 4 x 16): lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and holds D[l / 16 + 4 e][l % 16], e < 4.
__device__ __forceinline__ doublex4 mfma_f64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// C (M x N, ldc) = op(A) B over K, op(A)[m][k] = TA ? A[k * lda + m] : A[m * lda + k],
// B[k][n] = B[k * ldb + n]; M, N multiples of 16, K a multiple of 8.  The
// block's waves take 16 x 16 output tiles in turn; operands may be in LDS or
// global memory (L2-resident here).  Ends with a barrier.
template <bool TA>
__device__ void block_gemm_f64(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int M, int N,
                               int K) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const int tn = N / 16, ntile = (M / 16) * tn;
  for (int t = wave; t < ntile; t += blockDim.x >> 6) {
    const int mi = (t / tn) * 16, nj = (t % tn) * 16;
    doublex4 acc = doublex4{0.0, 0.0, 0.0, 0.0};
    // eight k-steps' operands requested before their MFMAs (k-ascending
    // accumulation either way)
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
      double a[8], b[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int kk = k0 + 4 * s + lk;
        a[s] = TA ? A[kk * lda + mi + lr] : A[(mi + lr) * lda + kk];
        b[s] = B[kk * ldb + nj + lr];
      }
#pragma unroll
      for (int s = 0; s < 8; ++s) acc = mfma_f64(a[s], b[s], acc);
    }
    for (; k0 < K; k0 += 4) {
      const int kk = k0 + lk;
      const double a0 = TA ? A[kk * lda + mi + lr] : A[(mi + lr) * lda + kk];
      acc = mfma_f64(a0, B[kk * ldb + nj + lr], acc);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) C[(mi + lk + 4 * e) * ldc + nj + lr] = acc[e];
  }
  __syncthreads();
}

// C (N x N) = A^T B for a product known to be symmetric (A^T A, or V^T G V
// with G symmetric): only the tiles on and above the diagonal are formed
// (16 k-steps' operands in flight) and each off-diagonal one is written to
// both places (a diagonal tile holds whatever rounding its own sums give).
// N a multiple of 16, K a multiple of 4.  Ends with a barrier.
__device__ void block_gemm_sym_f64(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int N,
                                   int K) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const int tn = N / 16, ntile = tn * (tn + 1) / 2;
  for (int t = wave; t < ntile; t += blockDim.x >> 6) {
    int ti = 0, rem = t;  // t -> (ti <= tj), row by row of the upper triangle
    while (rem >= tn - ti) {
      rem -= tn - ti;
      ++ti;
    }
    const int tj = ti + rem, mi = ti * 16, nj = tj * 16;
    doublex4 acc = doublex4{0.0, 0.0, 0.0, 0.0};
    int k0 = 0;
    for (; k0 + 64 <= K; k0 += 64) {
      double a[16], b[16];
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int kk = k0 + 4 * s + lk;
        a[s] = A[kk * lda + mi + lr];
        b[s] = B[kk * ldb + nj + lr];
      }
#pragma unroll
      for (int s = 0; s < 16; ++s) acc = mfma_f64(a[s], b[s], acc);
    }
    for (; k0 < K; k0 += 4) {
      const int kk = k0 + lk;
      acc = mfma_f64(A[kk * lda + mi + lr], B[kk * ldb + nj + lr], acc);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      C[(mi + lk + 4 * e) * ldc + nj + lr] = acc[e];
      if (ti != tj) C[(nj + lr) * ldc + mi + lk + 4 * e] = acc[e];
    }
  }
  __syncthreads();
}

// C (M x N) = A B for a row-major A in global memory with lda == K == KB:
// each wave stages one 16-row panel of A (16 * KB contiguous doubles, read
// with coalesced loads) in LDS and uses it for all N / 16 column tiles.
// Read straight from memory, an MFMA A operand is 16 rows x 32 bytes per
// load, which measured ~2.3x slower than contiguous pieces.  B (ldb) in LDS
// or global; the same k-ascending MFMA sequence as block_gemm_f64.  stage
// holds 16 * (KB + 1) doubles per wave.  Ends with a barrier.
template <int KB>
__device__ void block_gemm_f64_panel(const double* A, const double* B, int ldb, double* C, int ldc, int M, int N,
                                     double* stage) {
  constexpr int LDP = KB + 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  double* P = stage + wave * 16 * LDP;
  const int tn = N / 16;
  for (int pnl = wave; pnl < M / 16; pnl += blockDim.x >> 6) {
    const double* Ap = A + (size_t)pnl * 16 * KB;
#pragma unroll
    for (int e = lane; e < 16 * KB; e += 64) P[(e / KB) * LDP + (e % KB)] = Ap[e];
    // the wave's LDS writes reach LDS before its later reads (in-order LDS
    // queue per wave); this keeps the compiler from reordering across
    __builtin_amdgcn_wave_barrier();
    for (int nb = 0; nb < tn; ++nb) {
      doublex4 acc = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < KB; k0 += 4) {
        const int kk = k0 + lk;
        acc = mfma_f64(P[lr * LDP + kk], B[kk * ldb + nb * 16 + lr], acc);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) C[(pnl * 16 + lk + 4 * e) * ldc + nb * 16 + lr] = acc[e];
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
}

// Z = G V for the zero-padded symmetric np x np G and the np x B block V
// (both in global memory, ld np and B), read as G^T V so that each A-operand
// load takes 4 rows x 128 contiguous bytes of G.  V goes through LDS (vs,
// KC x B doubles) in KC-row chunks; each wave keeps two 16-row strips of Z
// with all B / 16 column tiles in registers, so one pass reads every element
// of G once and every V operand comes from LDS.  Same k-ascending MFMA
// sequence per output tile as block_gemm_f64.  Ends with a barrier.
template <int B, int KC>
__device__ void gv_product(const double* __restrict__ G, int np, const double* __restrict__ V, double* Z,
                           double* vs) {
  constexpr int TN = B / 16, SPW = 2;
  static_assert(B % 16 == 0 && KC % 32 == 0, "tile shape");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const int nw = blockDim.x >> 6, nstrip = np / 16;
  for (int g0 = 0; g0 < nstrip; g0 += nw * SPW) {
    const int s0 = g0 + wave * SPW;
    const bool live0 = s0 < nstrip, live1 = s0 + 1 < nstrip;  // wave-uniform
    doublex4 acc[SPW][TN];
#pragma unroll
    for (int s = 0; s < SPW; ++s)
#pragma unroll
      for (int t = 0; t < TN; ++t) acc[s][t] = doublex4{0.0, 0.0, 0.0, 0.0};
    for (int kc = 0; kc < np; kc += KC) {
      const int kn = min(KC, np - kc);  // a multiple of 16
      __syncthreads();                  // the previous chunk is consumed
      for (int e = threadIdx.x; e < kn * B; e += blockDim.x) vs[e] = V[(int64_t)kc * B + e];
      __syncthreads();
      if (live0) {
        const double* g0p = G + (int64_t)kc * np + s0 * 16 + lr;
        int k0 = 0;
        for (; k0 + 16 <= kn; k0 += 16) {
          double a0[4], a1[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int kk = k0 + 4 * q + lk;
            a0[q] = g0p[(int64_t)kk * np];
            a1[q] = live1 ? g0p[(int64_t)kk * np + 16] : 0.0;
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int kk = k0 + 4 * q + lk;
#pragma unroll
            for (int t = 0; t < TN; ++t) {
              const double bv = vs[kk * B + t * 16 + lr];
              acc[0][t] = mfma_f64(a0[q], bv, acc[0][t]);
              acc[1][t] = mfma_f64(a1[q], bv, acc[1][t]);
            }
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      const int strip = s0 + s;
      if (strip < nstrip) {
#pragma unroll
        for (int t = 0; t < TN; ++t)
#pragma unroll
          for (int e = 0; e < 4; ++e) Z[(int64_t)(strip * 16 + lk + 4 * e) * B + t * 16 + lr] = acc[s][t][e];
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Jacobi on an m x m symmetric matrix A (LDS, leading dimension LDA: odd, so
// that the column walks of a rotation hit distinct banks), accumulating the
// rotations into W (leading dimension LDW, in LDS or global memory,
// initialised by the caller).  Returns the number of sweeps.
// ---------------------------------------------------------------------------
// 1 / x and 1 / sqrt(x) for finite x > 0 (here x > 1e-300): the hardware
// estimates refined by two Newton steps each, to full double precision.
__device__ __forceinline__ double jac_recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double jac_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * fma(-hx * y, y, 1.5);
  return y * fma(-hx * y, y, 1.5);
}

// Slot a's pair (p < q) in round `round` of the round-robin ordering of me
// (even) slots: slot 0 stays, the others rotate.
__device__ __forceinline__ void jac_pair(int a, int round, int me, int& p, int& q) {
  const int L = me - 1;
  int x = (a == 0 ? me - 2 : a - 1) + round;
  if (x >= L) x -= L;
  int y = me - 2 - a + round;
  if (y >= L) y -= L;
  p = a == 0 ? 0 : 1 + x;
  q = a == 0 ? 1 + x : 1 + y;
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

// Items e = first, first + step, ... < count of an (outer, inner) grid with
// `inner` columns, walked without divisions: (o, i) advances by (dq, dr).
struct JacIter {
  int o, i, dq, dr, inner;
  __device__ JacIter(int first, int step, int inner_) : inner(inner_) {
    o = first / inner_;
    i = first - o * inner_;
    dq = step / inner_;
    dr = step - dq * inner_;
  }
  __device__ __forceinline__ void next() {
    o += dq;
    i += dr;
    if (i >= inner) {
      i -= inner;
      ++o;
    }
  }
};

// W <- W J for one round (cr: (c, s) per slot, pr: p | q << 16 per slot; a
// column index >= ma is the dummy slot): items (row < rows, slot) from
// `first` by `step` (it: the walk from `first`), four in flight.
template <int LDW>
__device__ __forceinline__ void jac_apply_w(double* W, const double2* cr, const int* pr, int rows, int ma, int npairs,
                                            int first, int step, JacIter it) {
  const int items = npairs * rows;
  for (int e0 = first; e0 < items; e0 += 4 * step) {
    double x[4], y[4];
    double2 cs[4];
    int p[4], q[4], row[4];
    bool on[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      on[u] = e0 + u * step < items;
      row[u] = it.o;
      const int pq = on[u] ? pr[it.i] : 0;
      p[u] = pq & 0xffff;
      q[u] = pq >> 16;
      on[u] = on[u] && q[u] < ma;  // the dummy slot has no column
      if (on[u]) {
        cs[u] = cr[it.i];
        x[u] = W[row[u] * LDW + p[u]];
        y[u] = W[row[u] * LDW + q[u]];
      }
      it.next();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (on[u]) {
        W[row[u] * LDW + p[u]] = cs[u].x * x[u] - cs[u].y * y[u];
        W[row[u] * LDW + q[u]] = cs[u].y * x[u] + cs[u].x * y[u];
      }
  }
}

// True when the `lead` largest eigenvalues of the symmetric A (m x m, LDS)
// belong to its leading rows: the Gershgorin discs of rows < lead lie above
// those of the other rows, and the coupling between the two sets is at most
// rel_off2 of ||A||_F^2 (uniform across the block; scratch[0..34]).
template <int LDA>
__device__ bool lead_block_separated(const double* A, int m, int lead, double rel_off2, double* scratch) {
  const int tid = threadIdx.x, nt = blockDim.x;
  // per row: Gershgorin radius, squared coupling to the other set, squares
  double lo = INFINITY, hi = -INFINITY, cpl = 0.0, tot = 0.0;
  if (tid < m) {
    double rad = 0.0;
    for (int j = 0; j < m; ++j) {
      const double v = A[tid * LDA + j];
      tot += v * v;
      if (j != tid) rad += fabs(v);
      if ((j < lead) != (tid < lead)) cpl += v * v;
    }
    const double d = A[tid * LDA + tid];
    if (tid < lead) lo = d - rad;
    else hi = d + rad;
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off, 64));
    hi = fmax(hi, __shfl_xor(hi, off, 64));
  }
  cpl = wave_sum(cpl);
  tot = wave_sum(tot);
  if ((tid & 63) == 0) {
    const int w = tid >> 6;
    scratch[w] = lo; scratch[8 + w] = hi; scratch[16 + w] = cpl; scratch[24 + w] = tot;
  }
  __syncthreads();
  if (tid == 0) {
    double l = INFINITY, h = -INFINITY, c = 0.0, t = 0.0;
    for (int w = 0; w < nt / 64 && w < 8; ++w) {
      l = fmin(l, scratch[w]); h = fmax(h, scratch[8 + w]); c += scratch[16 + w]; t += scratch[24 + w];
    }
    scratch[34] = (l > h && c <= rel_off2 * t) ? 1.0 : 0.0;
  }
  __syncthreads();
  const bool ok = scratch[34] != 0.0;
  __syncthreads();
  return ok;
}

// Eigenvectors of the leading ma x ma block of a symmetric A that is nearly
// diagonal, A = D + E (D its diagonal), by second-order perturbation theory
// instead of Jacobi sweeps: with M[k][i] = E[k][i] / (d_i - d_k) (k != i),
//   v_i = e_i + M e_i + sum_k e_k (E M)[k][i] / (d_i - d_k) - e_i sum_k M[k][i]^2 / 2,
//   lambda_i = d_i + sum_k E[k][i] M[k][i],
// correct to third order in max |M|.  Taken only when max |M| <= 1e-5 (no
// close eigenvalues relative to the coupling): the neglected terms, sums of
// up to ma^2 products of three entries of M, stay ~1e-12 of the vectors;
// else false and nothing is written.  On success W[0:ma, 0:ma] = the vectors
// (W is the identity on entry) and A's leading diagonal = the eigenvalues
// (its off-diagonal entries are left as they were).  Mbuf, Pbuf: ma x ma
// LDS scratch (stride ma).
template <int LDA, int LDW>
__device__ bool perturb_eigvecs(double* A, double* W, int ma, double* Mbuf, double* Pbuf, double* scratch) {
  const int tid = threadIdx.x, nt = blockDim.x;
  double mmax = 0.0;
  for (int e = tid; e < ma * ma; e += nt) {
    const int k = e / ma, i = e % ma;
    double mv = 0.0;
    if (k != i) mv = A[k * LDA + i] / (A[i * LDA + i] - A[k * LDA + k]);
    Mbuf[e] = mv;
    mmax = fmax(mmax, isfinite(mv) ? fabs(mv) : INFINITY);
  }
  for (int off = 32; off > 0; off >>= 1) mmax = fmax(mmax, __shfl_xor(mmax, off, 64));
  if ((tid & 63) == 0) scratch[tid >> 6] = mmax;
  __syncthreads();
  if (tid == 0) {
    double mx = 0.0;
    for (int w = 0; w < nt / 64 && w < 8; ++w) mx = fmax(mx, scratch[w]);
    scratch[35] = (mx <= 1e-5) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (scratch[35] == 0.0) {
    __syncthreads();
    return false;
  }
  // P = E M (E: A without its diagonal)
  for (int e = tid; e < ma * ma; e += nt) {
    const int k = e / ma, i = e % ma;
    double pv = 0.0;
    for (int l = 0; l < ma; ++l)
      if (l != k) pv = fma(A[k * LDA + l], Mbuf[l * ma + i], pv);
    Pbuf[e] = pv;
  }
  __syncthreads();
  for (int e = tid; e < ma * ma; e += nt) {
    const int k = e / ma, i = e % ma;
    double wv;
    if (k != i) {
      wv = Mbuf[e] + Pbuf[e] / (A[i * LDA + i] - A[k * LDA + k]);
    } else {
      double n2 = 0.0;
      for (int l = 0; l < ma; ++l) n2 = fma(Mbuf[l * ma + i], Mbuf[l * ma + i], n2);
      wv = 1.0 - 0.5 * n2;
    }
    W[k * LDW + i] = wv;
  }
  // eigenvalues after every thread has read the diagonal above
  double lam = 0.0;
  if (tid < ma) {
    lam = A[tid * LDA + tid];
    for (int k = 0; k < ma; ++k)
      if (k != tid) lam = fma(A[k * LDA + tid], Mbuf[k * ma + tid], lam);
  }
  __syncthreads();
  if (tid < ma) A[tid * LDA + tid] = lam;
  __syncthreads();
  return true;
}

// One sweep is me - 1 rounds of me / 2 disjoint rotations.  Per round: wave 0
// computes the round's rotations (c, s and the pair p | q << 16 per slot, in
// cs / pr, two rounds' worth each) while the other waves apply the previous
// round's to W; after a barrier every thread applies A <- J^T A J to whole
// 2 x 2 blocks (two in flight).  Every thread's items are fixed for the call
// and walked without divisions.
//
// lead > 0: only the eigenpairs of the `lead` largest eigenvalues are wanted.
// If the Gershgorin discs of rows < lead lie above those of the other rows
// (so the lead largest eigenvalues belong to the leading rows) and the
// coupling between the two sets is below the threshold
// (lead_block_separated; lead_known: its answer if already known, -1 not),
// the sweeps rotate the leading lead x lead block only; the coupling and the
// trailing block are left as they are (W must be the identity on entry; its
// trailing columns stay so).  Returns the number of sweeps.
template <int LDA, int LDW>
__device__ int lds_jacobi(double* A, double* W, int m, int max_sweeps, double2* cs, int* pr, double* scratch,
                          double rel_off2 = 1e-30, int lead = 0, int lead_known = -1) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int ma = m;  // the rows and columns the sweeps rotate
  if (lead > 0 && lead < m && (lead_known == 1 || (lead_known < 0 && lead_block_separated<LDA>(A, m, lead, rel_off2, scratch))))
    ma = lead;
  const int me = ma + (ma & 1);  // even number of slots (one dummy)
  const int npairs = me / 2, nblk = npairs * npairs;
  const JacIter blk0(tid, nt, npairs);                         // A blocks (pa, pb)
  const JacIter wside(tid < 64 ? 0 : tid - 64, nt - 64, npairs);  // W (row, slot), waves 1..
  const JacIter wall(tid, nt, npairs);                         // W (row, slot), every thread
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    // off-diagonal and total Frobenius norms of the rotated block
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < ma * ma; e += nt) {
      const int i = e / ma, j = e % ma;
      const double v = A[i * LDA + j];
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
    if (!(scratch[32] > rel_off2 * scratch[33]) || scratch[33] == 0.0) break;
    for (int round = 0; round < me - 1; ++round) {
      double2* cr = cs + (round & 1) * npairs;
      int* prr = pr + (round & 1) * npairs;
      if (tid < 64) {
        if (tid < npairs) {
          int p, q;
          jac_pair(tid, round, me, p, q);
          double c = 1.0, s = 0.0;
          if (q < ma) {
            const double apq = A[p * LDA + q], app = A[p * LDA + p], aqq = A[q * LDA + q];
            if (fabs(apq) > 1e-300) {
              // tan of the rotation angle t = sgn(tau) / (|tau| + sqrt(1 + tau^2)),
              // c = 1 / sqrt(1 + t^2); reciprocals and reciprocal square roots
              // from the hardware estimates refined by Newton steps (this
              // is on the critical path of every round)
              const double tau = (aqq - app) * jac_recip(2.0 * apq);
              const double at = fabs(tau);
              double t;
              if (at > 1e100) {
                t = 0.5 / at;
              } else {
                const double u = 1.0 + tau * tau;
                t = jac_recip(at + u * jac_rsqrt(u));
              }
              t = tau >= 0.0 ? t : -t;
              c = jac_rsqrt(1.0 + t * t);
              s = t * c;
            }
          }
          cr[tid] = double2{c, s};
          prr[tid] = p | (q << 16);
        }
      } else if (round > 0) {
        jac_apply_w<LDW>(W, cs + ((round - 1) & 1) * npairs, pr + ((round - 1) & 1) * npairs, ma, ma, npairs,
                         tid - 64, nt - 64, wside);
      }
      __syncthreads();
      // A <- J^T A J: each thread owns whole 2 x 2 blocks (rows of pair pa,
      // columns of pair pb), so no entry is touched twice
      JacIter it = blk0;
      for (int e0 = tid; e0 < nblk; e0 += 2 * nt) {
        double v[2][4];
        double2 ca[2], cb[2];
        int p0[2], q0[2], p1[2], q1[2];
        bool on[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          on[u] = e0 + u * nt < nblk;
          if (on[u]) {
            const int pqa = prr[it.o], pqb = prr[it.i];
            ca[u] = cr[it.o];
            cb[u] = cr[it.i];
            p0[u] = pqa & 0xffff; q0[u] = pqa >> 16;
            p1[u] = pqb & 0xffff; q1[u] = pqb >> 16;
            const bool r1 = q0[u] < ma, c1 = q1[u] < ma;  // the dummy slot has no row / column
            v[u][0] = A[p0[u] * LDA + p1[u]];
            v[u][1] = c1 ? A[p0[u] * LDA + q1[u]] : 0.0;
            v[u][2] = r1 ? A[q0[u] * LDA + p1[u]] : 0.0;
            v[u][3] = (r1 && c1) ? A[q0[u] * LDA + q1[u]] : 0.0;
          }
          it.next();
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          if (!on[u]) continue;
          const bool r1 = q0[u] < ma, c1 = q1[u] < ma;
          const double sa = ca[u].y, cc = ca[u].x, sb = cb[u].y, cd = cb[u].x;
          const double u00 = cc * v[u][0] - sa * v[u][2], u01 = cc * v[u][1] - sa * v[u][3];
          const double u10 = sa * v[u][0] + cc * v[u][2], u11 = sa * v[u][1] + cc * v[u][3];
          A[p0[u] * LDA + p1[u]] = cd * u00 - sb * u01;
          if (c1) A[p0[u] * LDA + q1[u]] = sb * u00 + cd * u01;
          if (r1) A[q0[u] * LDA + p1[u]] = cd * u10 - sb * u11;
          if (r1 && c1) A[q0[u] * LDA + q1[u]] = sb * u10 + cd * u11;
        }
      }
      __syncthreads();
    }
    // the last round's rotations into W
    jac_apply_w<LDW>(W, cs + ((me - 2) & 1) * npairs, pr + ((me - 2) & 1) * npairs, ma, ma, npairs, tid, nt, wall);
    __syncthreads();
  }
  return sweep;
}

// order[r] = index of the r-th largest diagonal entry of A (ties by index)
template <int LDA>
__device__ void sort_desc(const double* A, int m, int* order) {
  const int tid = threadIdx.x;
  if (tid < m) {
    const double v = A[tid * LDA + tid];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double u = A[j * LDA + j];
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
  int np;     // n rounded up to a multiple of 16 (subspace kernel)
  double* Gp;  // np x np zero-padded copy of G (subspace kernel)
  double* V;   // np x capacity, likewise Z, Y, U
  double* Z;
  double* Y;
  double* U;
  double* W;  // rotations when they do not fit LDS
  double* T;  // orthonormalising factor when it does not fit LDS
  // singular-vector outputs (tla_gram_left_singular; zero / NULL otherwise)
  int flip;        // apply the u-based sign rule to the k leading vectors
  int out_dtype;   // TLA_F32 / TLA_F64 for u_out and s_out
  void* u_out;     // n x k sign-corrected vectors in out_dtype
  void* s_out;     // k square roots of max(eigenvalue, 0) in out_dtype
  double* wblk;    // n x b: the whole final block (subspace kernel)
  int wblock;      // Q0 is a whole block (k0 == b) used without orthonormalising
};

// Output of one problem: the k leading eigenpairs are column colmap[r] (r
// without a map) of Ysrc (ld ldy) and theta[r].  evals / evecs as
// tla_sym_eig writes them (when non-NULL); with a.flip the reference's
// u-based sign rule first (each column multiplied by the sign of its entry
// of largest magnitude, the first such row on ties; sgn: k doubles of LDS);
// u_out / s_out in a.out_dtype; with a.wblk the first nb columns of Ysrc.
__device__ void eig_emit(const EigArgs& a, const double* Ysrc, int ldy, const int* colmap, const double* theta,
                         double* sgn, int nb) {
  const int tid = threadIdx.x, nt = blockDim.x, n = a.n, k = a.k;
  if (a.flip) {
    for (int r = tid >> 6; r < k; r += nt >> 6) {
      const int c = colmap ? colmap[r] : r;
      double best = -1.0;
      int bi = 0;
      for (int i = tid & 63; i < n; i += 64) {
        const double v = fabs(Ysrc[i * ldy + c]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if ((tid & 63) == 0) {
        const double y = Ysrc[bi * ldy + c];
        sgn[r] = y > 0.0 ? 1.0 : (y < 0.0 ? -1.0 : 0.0);
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < n * k; e += nt) {
    const int i = e / k, r = e - (e / k) * k;
    double v = Ysrc[i * ldy + (colmap ? colmap[r] : r)];
    if (a.flip) v *= sgn[r];
    if (a.evecs) a.evecs[e] = v;
    if (a.u_out) {
      if (a.out_dtype == TLA_F32) ((float*)a.u_out)[e] = (float)v;
      else ((double*)a.u_out)[e] = v;
    }
  }
  for (int r = tid; r < k; r += nt) {
    if (a.evals) a.evals[r] = theta[r];
    if (a.s_out) {
      const double sv = sqrt(fmax(theta[r], 0.0));
      if (a.out_dtype == TLA_F32) ((float*)a.s_out)[r] = (float)sv;
      else ((double*)a.s_out)[r] = sv;
    }
  }
  if (a.wblk)
    for (int e = tid; e < n * nb; e += nt) a.wblk[e] = Ysrc[(e / nb) * ldy + e % nb];
}

// Problems solved by one launch, one workgroup each (blockIdx.x selects).
constexpr int kEigMaxBatch = 8;
struct EigBatch {
  EigArgs a[kEigMaxBatch];
};

__device__ __forceinline__ double hash_unit(uint32_t i, uint32_t j) {
  uint32_t h = i * 0x9E3779B1u ^ (j + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
  return (double)h * (1.0 / 4294967296.0) - 0.5;
}

// Set-up of the subspace problems of one launch (blockIdx.y selects), spread
// over many workgroups: the zero-padded symmetric copy Gp = (G + G^T) / 2 in
// 32 x 32 tiles (the transposed tile of G read row by row through LDS), then
// the starting block Y (np x B): warm-start columns, then pseudo-random ones.
template <int B>
__global__ __launch_bounds__(256) void eig_prep_kernel(EigBatch batch, int gtiles) {
  const EigArgs a = batch.a[blockIdx.y];
  const int n = a.n, np = a.np, tid = threadIdx.x;
  if ((int)blockIdx.x < gtiles) {
    __shared__ double Tt[32][33];
    const int nt = (np + 31) / 32;
    if ((int)blockIdx.x >= nt * nt) return;
    const int ti = blockIdx.x / nt, tj = blockIdx.x % nt;
    for (int e = tid; e < 1024; e += 256) {  // Tt[r][c] = G[tj * 32 + r][ti * 32 + c]
      const int r = e >> 5, c = e & 31, gi = tj * 32 + r, gj = ti * 32 + c;
      Tt[r][c] = (gi < n && gj < n) ? a.G[(int64_t)gi * n + gj] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < 1024; e += 256) {
      const int r = e >> 5, c = e & 31, i = ti * 32 + r, j = tj * 32 + c;
      if (i < np && j < np) a.Gp[(int64_t)i * np + j] = (i < n && j < n) ? 0.5 * (a.G[(int64_t)i * n + j] + Tt[c][r]) : 0.0;
    }
    return;
  }
  const int64_t e = (int64_t)(blockIdx.x - gtiles) * 256 + tid;
  if (e >= (int64_t)np * B) return;
  const int i = (int)(e / B), j = (int)(e % B);
  a.Y[e] = i >= n ? 0.0 : ((a.Q0 && j < a.k0) ? a.Q0[(int64_t)i * a.q0_ld + j] : hash_unit(i, j));
}

// Jacobi on the whole matrix: CAP 64 keeps A and W in LDS; CAP 128 keeps A
// in LDS and W in the workspace (a.W, ld 128).
template <int CAP>
__global__ __launch_bounds__(kEigThreads) void sym_eig_small_kernel(EigBatch batch) {
  const EigArgs a = batch.a[blockIdx.x];
  constexpr int LDA = CAP + 1;
  __shared__ double A[CAP * LDA];
  __shared__ double Wl[CAP == kJac ? CAP * LDA : 1];
  __shared__ double2 cs[CAP + 2];
  __shared__ int pr[CAP + 2];
  __shared__ double scratch[64];
  __shared__ int order[CAP];
  __shared__ double theta[CAP];
  __shared__ double sgn[CAP];
  constexpr int LDW = CAP == kJac ? LDA : CAP;
  double* W = CAP == kJac ? Wl : a.W;
  const int n = a.n, tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    A[i * LDA + j] = 0.5 * (a.G[i * n + j] + a.G[j * n + i]);
    W[i * LDW + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int sweeps = lds_jacobi<LDA, LDW>(A, W, n, 60, cs, pr, scratch);
  sort_desc<LDA>(A, n, order);
  for (int r = tid; r < a.k; r += blockDim.x) theta[r] = A[order[r] * LDA + order[r]];
  __syncthreads();
  eig_emit(a, W, LDW, order, theta, sgn, 0);
  if (tid == 0 && a.iters) *a.iters = sweeps;
}

// Out = M T with T = L^-T D^-1/2 from S + shift I = L D L^T, S = M^T M: the
// Cholesky-QR orthonormalisation of the np x b block M (global, ld CAP),
// through a forward elimination of [S | I] held in registers by the block
// (TPR threads per row; per pivot its row and column go through LDS, one
// barrier).  S and the factor live in Sbuf (LDS, ld b + 1) and Tbuf (ld ldt).
//
// equil (dsc: B doubles of LDS): orthonormalise M D instead, D = diag(S)^-1/2
// (the same span; columns of very different lengths no longer spoil the
// conditioning), and report whether one pass suffices: true when every
// pivot of the equilibrated S is >= 1e-3 (cond(S) ~ 1e3 at most, so the
// columns come out orthonormal to ~1e-13).  Without equil it returns true.
template <int B>
__device__ bool block_orth(const double* M, double* Out, int np, double* Sbuf, double* Tbuf, int ldt, bool shift,
                           double* pbuf, double* dpiv, double* scratch, double* stage, bool equil = false,
                           double* dsc = nullptr) {
  constexpr int LDS_S = B + 1, CAP = B;
  // TPR threads per row of [S | I] (rows past B idle), CPT columns each
  constexpr int TPR = B > 64 ? 4 : 8, CPT = 2 * B / TPR;
  static_assert(TPR * B <= kEigThreads && CPT * TPR == 2 * B && TPR % 2 == 0, "thread layout");
  block_gemm_sym_f64(M, CAP, M, CAP, Sbuf, LDS_S, B, np);
  const int tid = threadIdx.x;
  if (equil) {
    if (tid < B) dsc[tid] = 1.0 / sqrt(fmax(Sbuf[tid * LDS_S + tid], 1e-300));
    __syncthreads();
    for (int e = tid; e < B * B; e += blockDim.x) {
      const int r = e / B, c = e % B;
      Sbuf[r * LDS_S + c] *= dsc[r] * dsc[c];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double mx = 0.0;
    for (int i = 0; i < B; ++i) mx = fmax(mx, Sbuf[i * LDS_S + i]);
    scratch[40] = mx;
  }
  __syncthreads();
  const double mx = scratch[40];
  const double sh = shift ? 1e-13 * mx * (double)np : 0.0;
  const double floor = 1e-300 + 1e-30 * mx;
  const int i = tid / TPR, seg = tid % TPR, c0 = seg * CPT;
  const bool live = i < B;
  double x[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = c0 + j;
    x[j] = !live ? 0.0 : (c < B ? Sbuf[i * LDS_S + c] + (c == i ? sh : 0.0) : ((c - B == i) ? 1.0 : 0.0));
  }
  for (int kb = 0; kb < TPR / 2; ++kb) {
#pragma unroll
    for (int kk = 0; kk < CPT; ++kk) {
      const int k = kb * CPT + kk;
      double* row = pbuf + (k & 1) * (3 * B);
      double* col = row + 2 * B;
      if (i == k) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) row[c0 + j] = x[j];
      }
      if (live && seg == kb) col[i] = x[kk];
      __syncthreads();
      const double p = fmax(col[k], floor);
      if (i == k && seg == 0) dpiv[k] = p;
      if (live && i > k) {
        const double f = col[i] / p;
#pragma unroll
        for (int j = 0; j < CPT; ++j) x[j] = fma(-f, row[c0 + j], x[j]);
      }
    }
  }
  if (equil && tid == 0) {
    double mn = INFINITY;
    for (int k = 0; k < B; ++k) mn = fmin(mn, dpiv[k]);
    scratch[41] = (mn >= 1e-3) ? 1.0 : 0.0;
  }
  __syncthreads();
  // x now holds row i of [D L^T | L^-1]; T[r][i] = L^-1[i][r] / sqrt(D[i])
  // (times dsc[r] when equilibrated)
  const double ds = live ? 1.0 / sqrt(dpiv[i]) : 0.0;
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int c = c0 + j;
    if (live && c >= B) Tbuf[(c - B) * ldt + i] = x[j] * ds * (equil ? dsc[c - B] : 1.0);
  }
  const bool well = !equil || scratch[41] != 0.0;
  __syncthreads();
  if constexpr (B == 48) block_gemm_f64_panel<B>(M, Tbuf, ldt, Out, CAP, np, B, stage);
  else block_gemm_f64<false>(M, CAP, Tbuf, ldt, Out, CAP, np, B, B);
  return well;
}

// Z = G V of the warm-block path spread over the GPU: one wave per 16 x 16
// tile of Z (blockIdx.x: tile, blockIdx.y: problem), V = Y (the warm block,
// src 0) or V (src 1).  G read as G^T, the same k-ascending MFMA sequence per
// tile as gv_product, so Z is the same to the bit.
template <int B>
__global__ __launch_bounds__(64) void eig_gv_kernel(EigBatch batch, int src) {
  const EigArgs a = batch.a[blockIdx.y];
  constexpr int TN = B / 16;
  const int np = a.np;
  if ((int)blockIdx.x >= (np / 16) * TN) return;
  const int mi = (blockIdx.x / TN) * 16, nj = (blockIdx.x % TN) * 16;
  const double* V = src ? a.V : a.Y;
  const int lane = threadIdx.x, lr = lane & 15, lk = lane >> 4;
  const double* gp = a.Gp + mi + lr;
  doublex4 acc = doublex4{0.0, 0.0, 0.0, 0.0};
  int k0 = 0;
  for (; k0 + 64 <= np; k0 += 64) {
    double x[16], y[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int kk = k0 + 4 * s + lk;
      x[s] = gp[(int64_t)kk * np];
      y[s] = V[kk * B + nj + lr];
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) acc = mfma_f64(x[s], y[s], acc);
  }
  for (; k0 < np; k0 += 4) {
    const int kk = k0 + lk;
    acc = mfma_f64(gp[(int64_t)kk * np], V[kk * B + nj + lr], acc);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) a.Z[(int64_t)(mi + lk + 4 * e) * B + nj + lr] = acc[e];
}

// Block subspace iteration with Rayleigh-Ritz for n > kJacLarge.  B is the
// block: 48 and 64 keep H, W and T in LDS; 128 keeps H in LDS and W, T in the
// workspace.  All products are FP64 MFMA on zero-padded np-row blocks.
// phase 0: the whole solve.  A warm block (a.wblock) instead runs as
// eig_gv_kernel (Z = G Y), phase 1 (V <- orth(Z)), eig_gv_kernel (Z = G V),
// phase 2 (the solve from its first Rayleigh-Ritz step), so that both
// products spread over the GPU.
template <int B>
__global__ __launch_bounds__(kEigThreads) void sym_eig_subspace_kernel(EigBatch batch, int phase) {
  const EigArgs a = batch.a[blockIdx.x];
  constexpr int LDH = B + 1, CAP = B;
  constexpr bool SMALL = B <= 64;
  __shared__ double H[B * LDH];  // Ritz matrix; also S of the orthonormalisation; then sorted rotations
  __shared__ double Wl[SMALL ? B * LDH : 1];
  __shared__ double Tl[SMALL ? B * LDH : 1];
  __shared__ double2 cs[B + 2];
  __shared__ int pr[B + 2];
  __shared__ double scratch[64];
  __shared__ double resid[B];
  __shared__ int order[B];
  __shared__ double theta[B];
  __shared__ double pbuf[6 * B];
  __shared__ double dpiv[B];
  // 16-row panels of the np x B blocks for the products with B x B factors
  __shared__ double stage[B == 48 ? kEigWaves * 16 * (B + 1) : 1];
  constexpr int LDW = SMALL ? LDH : CAP;
  double* W = SMALL ? Wl : a.W;
  double* T = SMALL ? Tl : a.T;
  const int ldt = SMALL ? LDH : CAP;
  const int n = a.n, np = a.np, k = a.k, tid = threadIdx.x, nt = blockDim.x;
  double* V = a.V;
  double* Z = a.Z;
  double* Y = a.Y;
  double* U = a.U;
  // Gp (zero-padded symmetric G) and the starting block Y come from
  // eig_prep_kernel, launched just before on the same stream
  if (phase == 1) {
    if (!block_orth<B>(Z, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage, true, resid)) {
      block_orth<B>(Z, U, np, H, T, ldt, true, pbuf, dpiv, scratch, stage);
      block_orth<B>(U, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage);
    }
    return;
  }
  const double jfloor = fmax(1e-30, 1e-2 * a.tol * a.tol);
  double jthr = 1e-10;
  int it = 0;
  bool have_z = false;  // phase 2: Z = G V is there already
  if (phase == 2) {
    jthr = jfloor;
    it = 1;
    have_z = true;
  } else if (a.Q0 && a.k0 > 0 && a.wblock) {
    // a whole block from an earlier solve: the plain subspace step below
    // orthonormalises G V, so V itself is used as it is
    double* t = V;
    V = Y;
    Y = t;
  } else if (a.Q0 && a.k0 > 0) {
    // orthonormal warm-start columns and B - k0 pseudo-random ones: the
    // block is well conditioned (cond(S) ~ 1e2), one unshifted Cholesky-QR
    // pass leaves V orthonormal to ~1e-13
    block_orth<B>(Y, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage);
  } else {
    block_orth<B>(Y, V, np, H, T, ldt, true, pbuf, dpiv, scratch, stage);
    block_orth<B>(V, Y, np, H, T, ldt, false, pbuf, dpiv, scratch, stage);
    double* t = V;
    V = Y;
    Y = t;
  }
  if (phase == 0 && a.Q0 && a.k0 > 0 && a.max_iter >= 2) {
    // warm start: its first Ritz test fails whenever the matrix has moved by
    // more than the tolerance, so the first step is a plain subspace step,
    // V <- orth(G V) (the span a Rayleigh-Ritz step would leave), and the
    // next Rayleigh-Ritz step rotates to the final threshold
    if constexpr (B == 48) gv_product<B, 128>(a.Gp, np, V, Z, stage);
    else block_gemm_f64<true>(a.Gp, np, V, CAP, Z, CAP, np, B, np);
    // G V's columns are close to multiples of the warm columns: once
    // equilibrated, one unshifted Cholesky-QR pass usually suffices
    if (!block_orth<B>(Z, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage, true, resid)) {
      block_orth<B>(Z, U, np, H, T, ldt, true, pbuf, dpiv, scratch, stage);
      block_orth<B>(U, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage);
    }
    jthr = jfloor;
    it = 1;
  } else if (phase == 0 && !(a.Q0 && a.k0 > 0)) {
    // cold start: two plain subspace steps V <- orth(G V) before the first
    // Rayleigh-Ritz step (whose Ritz matrix on a random block is dense and
    // costs full Jacobi sweeps); a plain step is a fraction of a Ritz step
    for (int p = 0; p < 2 && it + 2 < a.max_iter; ++p, ++it) {
      if constexpr (B == 48) gv_product<B, 128>(a.Gp, np, V, Z, stage);
      else block_gemm_f64<true>(a.Gp, np, V, CAP, Z, CAP, np, B, np);
      if (!block_orth<B>(Z, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage, true, resid)) {
        block_orth<B>(Z, U, np, H, T, ldt, true, pbuf, dpiv, scratch, stage);
        block_orth<B>(U, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage);
      }
    }
  }
  for (; it < a.max_iter; ++it) {
    // Z = G V, read as G^T V (Gp is exactly symmetric): each load then takes
    // 4 rows x 128 contiguous bytes of G instead of 16 rows x 32 bytes, which
    // measured 2.3x faster
    if (!have_z) {
      if constexpr (B == 48) gv_product<B, 128>(a.Gp, np, V, Z, stage);
      else block_gemm_f64<true>(a.Gp, np, V, CAP, Z, CAP, np, B, np);
    }
    have_z = false;
    block_gemm_sym_f64(V, CAP, Z, CAP, H, LDH, B, np);           // H = V^T Z = V^T G V
    for (int e = tid; e < B * B; e += nt) {
      const int i = e / B, j = e % B;
      if (i < j) H[i * LDH + j] = H[j * LDH + i] = 0.5 * (H[i * LDH + j] + H[j * LDH + i]);
      W[i * LDW + j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    // the Ritz rotations only need the accuracy of the current subspace:
    // off-diagonal mass down to (residual / t_1)^2 / 100 of the total,
    // tightening to tol^2 / 100 as the residuals fall
    {
      // a nearly diagonal leading block (a converging warm start): second-
      // order perturbation instead of Jacobi sweeps when it is safe (block
      // 48: its two k x k scratch matrices use the panel staging area)
      const bool sep = k < B && lead_block_separated<LDH>(H, B, k, jthr, scratch);
      bool pert = false;
      if constexpr (B == 48) {
        if (sep) pert = perturb_eigvecs<LDH, LDW>(H, W, k, stage, stage + k * k, scratch);
      }
      if (!pert) lds_jacobi<LDH, LDW>(H, W, B, 30, cs, pr, scratch, jthr, k, sep ? 1 : 0);
    }
    sort_desc<LDH>(H, B, order);
    if (tid < B) theta[tid] = H[order[tid] * LDH + order[tid]];
    __syncthreads();
    for (int e = tid; e < B * B; e += nt) {  // rotations in Ritz-value order
      const int l = e / B, j = e % B;
      H[l * LDH + j] = W[l * LDW + order[j]];
    }
    __syncthreads();
    if constexpr (B == 48) {
      block_gemm_f64_panel<B>(V, H, LDH, Y, CAP, np, B, stage);  // Y = V W (Ritz vectors)
      block_gemm_f64_panel<B>(Z, H, LDH, U, CAP, np, B, stage);  // U = Z W = G Y
    } else {
      block_gemm_f64<false>(V, CAP, H, LDH, Y, CAP, np, B, B);
      block_gemm_f64<false>(Z, CAP, H, LDH, U, CAP, np, B, B);
    }
    // residual norms of the leading k pairs: thread (g, r) sums column r
    // over rows g, g + ng, ... (row-wise, coalesced), partials through H
    {
      const int ng = nt / k;
      if (tid < ng * k) {
        const int g = tid / k, r = tid - g * k;
        const double th = theta[r];
        double r2 = 0.0;
        for (int i = g; i < n; i += ng) {
          const double d = U[i * CAP + r] - th * Y[i * CAP + r];
          r2 += d * d;
        }
        H[tid] = r2;
      }
      __syncthreads();
      if (tid < k) {
        double r2 = 0.0;
        for (int g = 0; g < ng; ++g) r2 += H[g * k + tid];
        resid[tid] = sqrt(r2);
      }
    }
    __syncthreads();
    if (tid == 0) {
      double worst = 0.0;
      for (int r = 0; r < k; ++r) worst = fmax(worst, resid[r]);
      const double t1 = fmax(fabs(theta[0]), 1e-300);
      scratch[63] = (worst <= a.tol * t1) ? 1.0 : 0.0;
      const double rr = worst / t1;
      scratch[62] = fmin(1e-10, fmax(jfloor, 1e-2 * rr * rr));
    }
    __syncthreads();
    const bool done = scratch[63] != 0.0;
    jthr = scratch[62];
    __syncthreads();
    if (done || it + 1 == a.max_iter) break;
    // V <- orth(U): the columns of U = G Y scale with the Ritz values, so
    // equilibrated first, one pass when that leaves S well conditioned,
    // else two Cholesky-QR passes (shifted, then plain)
    if (!block_orth<B>(U, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage, true, resid)) {
      block_orth<B>(U, Z, np, H, T, ldt, true, pbuf, dpiv, scratch, stage);
      block_orth<B>(Z, V, np, H, T, ldt, false, pbuf, dpiv, scratch, stage);
    }
  }
  eig_emit(a, Y, CAP, nullptr, theta, resid, B);
  if (tid == 0 && a.iters) *a.iters = it + 1;
}

// ---------------------------------------------------------------------------
// Mode Gram: G[i, j] = sum_{a,b} X[a, i, b] X[a, j, b], FP64 accumulation.
// Block (ti, tj) computes a 32 x 32 tile for one chunk of the (a, b) pairs
// (256 threads: four waves, one 16 x 16 FP64 MFMA tile each); partial tiles
// are summed in chunk order.
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
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int mi = (wave >> 1) * 16, nj = (wave & 1) * 16;
  doublex4 acc = doublex4{0.0, 0.0, 0.0, 0.0};
  // thread: column qq of the 64-wide chunk, rows r0, r0 + 4, ..., r0 + 28 of
  // both 32-row blocks; the chunk after the current one is loaded into
  // registers while the current one is multiplied
  const int qq = tid & 63, r0 = tid >> 6;
  T ci[8], cj[8];
  auto fetch = [&](int64_t qb) {
    const int64_t q = qb + qq;
    int64_t base = -1;
    if (q < q1) {
      const int64_t aa = q / Pb;
      base = aa * In * Pb + (q - aa * Pb);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t gi = (int64_t)ti * 32 + r0 + 4 * u, gj = (int64_t)tj * 32 + r0 + 4 * u;
      ci[u] = (base >= 0 && gi < In) ? X[base + gi * Pb] : (T)0;
      cj[u] = (base >= 0 && gj < In) ? X[base + gj * Pb] : (T)0;
    }
  };
  if (q0 < q1) fetch(q0);
  for (int64_t qb = q0; qb < q1; qb += 64) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      Ti[r0 + 4 * u][qq] = (double)ci[u];
      Tj[r0 + 4 * u][qq] = (double)cj[u];
    }
    __syncthreads();
    if (qb + 64 < q1) fetch(qb + 64);
    // wave w: the 16 x 16 output tile (w / 2, w % 2) of the 32 x 32 block on
    // FP64 MFMA (products of the float inputs are exact in FP64)
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ti[mi + lr][k0 + lk], Tj[nj + lr][k0 + lk], acc, 0, 0, 0);
    __syncthreads();
  }
  double* out = partial + (int64_t)blockIdx.z * In * In;
  // lane holds D[lk + 4 e][lr]
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int64_t i = (int64_t)ti * 32 + mi + lk + 4 * e, j = (int64_t)tj * 32 + nj + lr;
    if (i < In && j < In) {
      out[i * In + j] = acc[e];
      out[j * In + i] = acc[e];
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

static int64_t eig_np(int n) { return (int64_t)(n + 15) / 16 * 16; }

extern "C" size_t tla_sym_eig_workspace_size(int n) {
  if (n <= kJac) return 0;
  const size_t rot = align256(size_t(kEigCapLarge) * kEigCapLarge * 8);
  if (n <= kJacLarge) return rot;  // the rotations of the Jacobi on the whole matrix
  // padded G, four np x 128 blocks, rotations and the orthonormalising factor
  const size_t np = (size_t)eig_np(n);
  return align256(np * np * 8) + 4 * align256(np * kEigCapLarge * 8) + 2 * rot;
}

// Arguments of one problem (workspace carved as the kernel kind needs);
// *kind: 0 Jacobi in LDS, 1 Jacobi with rotations in the workspace, 2/3/4
// subspace iteration with block 48 / 64 / 128.
static int eig_args(const double* G, int n, int k, const double* Q0, int k0, double* evals, double* evecs, double tol,
                    int max_iter, int* iters_out, void* ws, size_t ws_bytes, EigArgs* out, int* kind) {
  TLA_REQUIRE(n >= 1, "sym_eig: n = %d", n);
  TLA_REQUIRE(k >= 1 && k <= n, "sym_eig: k = %d not in [1, n = %d]", k, n);
  TLA_REQUIRE(G, "sym_eig: NULL pointer");
  TLA_REQUIRE(!Q0 || k0 >= 1, "sym_eig: warm start with k0 = %d columns", k0);
  TLA_REQUIRE(ws_bytes >= tla_sym_eig_workspace_size(n), "sym_eig: workspace too small");
  TLA_REQUIRE(n <= kJacLarge || ws, "sym_eig: NULL workspace");
  EigArgs a{};
  a.G = G; a.n = n; a.k = k; a.Q0 = Q0; a.q0_ld = k0;
  a.evals = evals; a.evecs = evecs; a.tol = tol > 0 ? tol : 1e-11; a.max_iter = max_iter > 0 ? max_iter : 200;
  a.iters = iters_out;
  if (n <= kJac) {
    *out = a;
    *kind = 0;
    return TLA_OK;
  }
  if (n <= kJacLarge) {
    a.W = (double*)ws;
    *out = a;
    *kind = 1;
    return TLA_OK;
  }
  // block: 48 for k <= 40 (cheaper Jacobi sweeps), 64 for k <= 56, 128 for k <= 120
  const bool small_block = k <= 56;
  if (!small_block && k > kEigCapLarge - 8) {
    set_error("sym_eig: k = %d > %d leading eigenpairs of an n = %d > %d matrix is not supported", k,
              kEigCapLarge - 8, n, kJacLarge);
    return TLA_ERR_UNSUPPORTED;
  }
  const int b = k <= 40 ? 48 : (small_block ? 64 : kEigCapLarge);
  a.np = (int)eig_np(n);
  const size_t np = (size_t)a.np;
  const size_t blk = align256(np * kEigCapLarge * 8);
  const size_t rot = align256(size_t(kEigCapLarge) * kEigCapLarge * 8);
  char* base = (char*)ws;
  a.Gp = (double*)base;
  base += align256(np * np * 8);
  a.V = (double*)base;
  a.Z = (double*)(base + blk);
  a.Y = (double*)(base + 2 * blk);
  a.U = (double*)(base + 3 * blk);
  a.W = (double*)(base + 4 * blk);
  a.T = (double*)(base + 4 * blk + rot);
  a.k0 = Q0 ? std::min(k0, b) : 0;
  *out = a;
  *kind = b == 48 ? 2 : (b == 64 ? 3 : 4);
  return TLA_OK;
}

// The warm-block sequence of sym_eig_subspace_kernel (m problems).
template <int B>
static void launch_split(const EigBatch& b, int m, hipStream_t st) {
  int tiles = 0;
  for (int i = 0; i < m; ++i) tiles = std::max(tiles, (b.a[i].np / 16) * (B / 16));
  const dim3 g((unsigned)tiles, (unsigned)m);
  eig_gv_kernel<B><<<g, 64, 0, st>>>(b, 0);
  sym_eig_subspace_kernel<B><<<m, kEigThreads, 0, st>>>(b, 1);
  eig_gv_kernel<B><<<g, 64, 0, st>>>(b, 1);
  sym_eig_subspace_kernel<B><<<m, kEigThreads, 0, st>>>(b, 2);
}

// One launch per kernel kind present, one workgroup per problem.
static int eig_launch(const EigArgs* args, const int* kinds, int count, hipStream_t st) {
  for (int kind = 0; kind < 5; ++kind) {
    EigBatch b{};
    int m = 0;
    for (int i = 0; i < count; ++i)
      if (kinds[i] == kind) b.a[m++] = args[i];
    if (m == 0) continue;
    if (kind >= 2) {
      int gt = 0, yt = 0;
      const int bk = kind == 2 ? 48 : (kind == 3 ? 64 : kEigCapLarge);
      for (int i = 0; i < m; ++i) {
        const int nt = (b.a[i].np + 31) / 32;
        gt = std::max(gt, nt * nt);
        yt = std::max(yt, (int)ceil_div((int64_t)b.a[i].np * bk, 256));
      }
      dim3 grid(gt + yt, m);
      switch (kind) {
        case 2: eig_prep_kernel<48><<<grid, 256, 0, st>>>(b, gt); break;
        case 3: eig_prep_kernel<64><<<grid, 256, 0, st>>>(b, gt); break;
        default: eig_prep_kernel<kEigCapLarge><<<grid, 256, 0, st>>>(b, gt); break;
      }
      TLA_LAUNCH_CHECK();
    }
    bool split = kind >= 2;  // every problem a warm block
    for (int i = 0; i < m; ++i) split = split && b.a[i].wblock;
    switch (kind) {
      case 0: sym_eig_small_kernel<kJac><<<m, kEigThreads, 0, st>>>(b); break;
      case 1: sym_eig_small_kernel<kJacLarge><<<m, kEigThreads, 0, st>>>(b); break;
      case 2: split ? launch_split<48>(b, m, st) : sym_eig_subspace_kernel<48><<<m, kEigThreads, 0, st>>>(b, 0); break;
      case 3: split ? launch_split<64>(b, m, st) : sym_eig_subspace_kernel<64><<<m, kEigThreads, 0, st>>>(b, 0); break;
      default: sym_eig_subspace_kernel<kEigCapLarge><<<m, kEigThreads, 0, st>>>(b, 0); break;
    }
    TLA_LAUNCH_CHECK();
  }
  return TLA_OK;
}

extern "C" int tla_sym_eig(const double* G, int n, int k, const double* Q0, int k0, double* evals, double* evecs,
                           double tol, int max_iter, int* iters_out, void* ws, size_t ws_bytes, void* stream) {
  TLA_REQUIRE(evals && evecs, "sym_eig: NULL pointer");
  EigArgs a;
  int kind = 0;
  const int rc = eig_args(G, n, k, Q0, k0, evals, evecs, tol, max_iter, iters_out, ws, ws_bytes, &a, &kind);
  if (rc != TLA_OK) return rc;
  return eig_launch(&a, &kind, 1, (hipStream_t)stream);
}

extern "C" int tla_gram_svd_block(int n, int k) {
  if (n <= kJacLarge || k < 1 || k > n || k > kEigCapLarge - 8) return 0;
  return k <= 40 ? 48 : (k <= 56 ? 64 : kEigCapLarge);
}

extern "C" int tla_gram_left_singular(const double* G, int n, int k, const double* W0, int k0, int dtype, void* U,
                                      void* s, double* U64, double* W_out, double tol, int max_iter, int* iters_out,
                                      void* ws, size_t ws_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "gram_left_singular: bad dtype %d", dtype);
  TLA_REQUIRE(U, "gram_left_singular: NULL U");
  EigArgs a;
  int kind = 0;
  const int rc = eig_args(G, n, k, W0, W0 ? k0 : 0, nullptr, U64, tol, max_iter, iters_out, ws, ws_bytes, &a, &kind);
  if (rc != TLA_OK) return rc;
  const int b = tla_gram_svd_block(n, k);
  a.flip = 1;
  a.out_dtype = dtype;
  a.u_out = U;
  a.s_out = s;
  a.wblk = b > 0 ? W_out : nullptr;
  a.wblock = (b > 0 && W0 && k0 == b && a.max_iter >= 2) ? 1 : 0;
  return eig_launch(&a, &kind, 1, (hipStream_t)stream);
}

extern "C" int tla_sym_eig_batch(int count, const double* const* G, const int* n, const int* k,
                                 const double* const* Q0, const int* k0, double* const* evals, double* const* evecs,
                                 double tol, int max_iter, void* const* workspace, const size_t* workspace_bytes,
                                 void* stream) {
  TLA_REQUIRE(count >= 1 && count <= kEigMaxBatch, "sym_eig_batch: count %d not in [1, %d]", count, kEigMaxBatch);
  TLA_REQUIRE(G && n && k && evals && evecs && workspace && workspace_bytes, "sym_eig_batch: NULL array");
  EigArgs args[kEigMaxBatch];
  int kinds[kEigMaxBatch];
  for (int i = 0; i < count; ++i) {
    TLA_REQUIRE(evals[i] && evecs[i], "sym_eig_batch: NULL pointer");
    const int rc = eig_args(G[i], n[i], k[i], Q0 ? Q0[i] : nullptr, k0 ? k0[i] : 0, evals[i], evecs[i], tol,
                            max_iter, nullptr, workspace[i], workspace_bytes[i], &args[i], &kinds[i]);
    if (rc != TLA_OK) return rc;
  }
  return eig_launch(args, kinds, count, (hipStream_t)stream);
}
