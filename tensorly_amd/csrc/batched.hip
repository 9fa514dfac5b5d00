// Batched CP-ALS: thousands of small dense 3-way problems (mode sizes <= 64,
// rank <= 16) solved to convergence in one persistent launch.
//
// Each workgroup takes the next problem from a device work queue (a vector
// atomic on one counter; the host orders the queue by decreasing estimated
// cost) and runs the whole ALS loop of tensorly/decomposition/_cp.py:394-531
// for it:
//   * factors live in LDS (row-major, rank padded to 16 columns with zeros);
//   * MTTKRP of mode n is an MFMA product (v_mfma_f32_16x16x4_f32) of the
//     tensor rows against Khatri-Rao rows formed on the fly from the two
//     other factors in LDS; the reduction index is split over the waves and
//     the wave partials are summed in a fixed order (deterministic);
//   * V = Hadamard of the FP64 Grams, inverted by Gauss-Jordan with partial
//     pivoting in FP64 (the LU solve of _cp.py:425-427), A_n = M V^-1;
//   * the relative error of _cp.py:217-225 from the Grams in FP64.
// Policies per launch: plain ALS, the reference line search
// (_cp.py:333-336,401-403,448-483: jump = iteration^(1/acc_pow) every second
// iteration from 6, judged by the error of the extrapolated point), the
// learned step (16 -> 64 -> 64 -> 1 network on the error history, kept only
// if the sweep run from the extrapolated point lowers the error), and a
// probing policy that records, at every decision, the error reached after one
// sweep for each step of a grid (training data for the network).
#include <math.h>

#include "common.h"
#include "extrap.h"

namespace tla {
namespace {

constexpr int kBMaxDim = 64;
constexpr int kBMaxRank = 16;
constexpr int kBTiles = kBMaxDim / 16;
constexpr int kBFac = kBMaxDim * 16;  // floats per factor in LDS
constexpr int kMaxProbe = 16;

enum { kPlain = 0, kRefLineSearch = 1, kLearned = 2, kProbe = 3 };
enum { kStOk = 0, kStSingular = 2, kStNonFinite = 3 };

struct BatchArgs {
  int P;
  const float* X;
  const int64_t* x_off;
  const int* shape;  // [P][3]
  const int* rank;   // [P]
  float* factors;    // per problem: A (I0 x R) | B (I1 x R) | C (I2 x R), row-major
  const int64_t* f_off;
  float* f_prev;  // scratch, same layout as factors
  float* f_save;  // scratch, same layout as factors
  const int* order;
  int* queue;
  int policy;
  int max_sweeps;
  double tol;
  int warmup;
  const float* mlp;
  float alpha_fixed, alpha_max;
  const float* probe_alphas;
  int n_probe;
  int max_dec;
  int first_dec;  // probing covers decisions first_dec .. first_dec + max_dec - 1
  double* hist_err;  // [P][max_sweeps]
  float* hist_alpha;
  int* hist_acc;
  float* feats;       // [P][max_dec][16]
  double* probe_err;  // [P][max_dec][n_probe]
  int* sweeps_out;
  int* passes_out;
  double* err_out;
  int* n_extrap_out;
  int* n_accept_out;
  int* status_out;
  long long* ticks_out;
  int kdepth;  // reduction tile of the MTTKRP: 16 * kdepth indices in flight per wave (1, 2 or 4)
};

struct Prob {
  const float* X;
  float* fac;
  float* prev;
  float* save;
  int I0, I1, I2;
  int R;
  int64_t N;
  // selects instead of arrays indexed by a runtime mode (those would live in scratch)
  __device__ int dim(int m) const { return m == 0 ? I0 : (m == 1 ? I1 : I2); }
  __device__ int foff(int m) const { return m == 0 ? 0 : (m == 1 ? I0 * R : (I0 + I1) * R); }
};

template <int NW>
struct Smem {
  Prob prob;
  float F[3][kBFac];
  float part[NW][kBTiles][256];  // wave partials of the MTTKRP; part[0] then holds M (a * 16 + r)
  double G[3][256];
  double aug[kBMaxRank][2 * kBMaxRank + 1];
  double red[NW];
  float feat[kMlpIn];
  float h1[kMlpH];
  float hpart[kMlpH];
  double dscal[4];
  int iscal[4];
  int kdepth;
};

// One instance per workgroup (LDS); every helper below refers to it directly,
// so the compiler keeps LDS addressing in the non-inlined helpers too.
template <int NW>
__shared__ Smem<NW> g_sm;


template <int NW>
__device__ inline double block_sum(double v) {
  Smem<NW>& sm = g_sm<NW>;
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sm.red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NW; ++i) s += sm.red[i];
  return s;
}

// Factor block (global, row-major I_m x R) -> LDS (rows padded to 64, columns to 16).
template <int NW>
__device__ void load_factors(const float* src) {
  Smem<NW>& sm = g_sm<NW>;
  const Prob& p = sm.prob;
  for (int m = 0; m < 3; ++m) {
    for (int i = threadIdx.x; i < kBFac; i += NW * 64) {
      const int a = i >> 4, r = i & 15;
      sm.F[m][i] = (a < p.dim(m) && r < p.R) ? src[p.foff(m) + a * p.R + r] : 0.0f;
    }
  }
  __syncthreads();
}

template <int NW>
__device__ void store_factors(float* dst) {
  Smem<NW>& sm = g_sm<NW>;
  const Prob& p = sm.prob;
  for (int m = 0; m < 3; ++m) {
    const int n = p.dim(m) * p.R;
    for (int i = threadIdx.x; i < n; i += NW * 64) {
      const int a = i / p.R, r = i - a * p.R;
      dst[p.foff(m) + i] = sm.F[m][a * 16 + r];
    }
  }
  __syncthreads();
}

// F <- F + alpha (F - base), base a factor block in global memory.
template <int NW>
__device__ void extrapolate(const float* base, float alpha) {
  Smem<NW>& sm = g_sm<NW>;
  const Prob& p = sm.prob;
  for (int m = 0; m < 3; ++m) {
    const int n = p.dim(m) * p.R;
    for (int i = threadIdx.x; i < n; i += NW * 64) {
      const int a = i / p.R, r = i - a * p.R;
      const float x = sm.F[m][a * 16 + r];
      sm.F[m][a * 16 + r] = x + alpha * (x - base[p.foff(m) + i]);
    }
  }
  __syncthreads();
}

template <int NW>
__device__ void gram(int m) {
  Smem<NW>& sm = g_sm<NW>;
  const Prob& p = sm.prob;
  for (int idx = threadIdx.x; idx < 256; idx += NW * 64) {
    const int r = idx >> 4, s = idx & 15;
    double acc = 0.0;
    const float* f = sm.F[m];
    for (int a = 0; a < p.dim(m); ++a) acc += (double)f[a * 16 + r] * (double)f[a * 16 + s];
    sm.G[m][idx] = acc;
  }
  __syncthreads();
}

template <int NW>
__device__ void grams_all() {
  for (int m = 0; m < 3; ++m) gram<NW>(m);
}

// M (part[0], a * 16 + r) = unfold(X, n) . KR(F_o, F_i) on MFMA, where o < i
// are the two other modes.  Reduction index e = (index along o) * I_i +
// (index along i); wave w takes a contiguous range of e.  The reduction
// tile: D chunks of 16 reduction indices are in flight per wave (the loads of
// chunk c + D are issued after the MFMAs of chunk c); every D gives the same
// MFMA sequence, so the same bits.
template <int NW, int T, int D>
__device__ __forceinline__ void mttkrp_tiles(int n) {
  Smem<NW>& sm = g_sm<NW>;
  const Prob& p = sm.prob;
  const int mo = (n == 0) ? 1 : 0;
  const int mi = (n == 2) ? 1 : 2;
  const int s01 = p.I1 * p.I2;
  const int s_a = n == 0 ? s01 : (n == 1 ? p.I2 : 1);
  const int s_o = mo == 0 ? s01 : p.I2;  // mo is 0 or 1
  const int s_i = mi == 2 ? 1 : p.I2;    // mi is 1 or 2
  const int n_in = p.dim(mi);
  const int E = p.dim(mo) * n_in;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int chunk = ((E + NW * 4 - 1) / (NW * 4)) * 4;
  const int e0 = min(E, w * chunk), e1 = min(E, e0 + chunk);
  const float* Fo = sm.F[mo];
  const float* Fi = sm.F[mi];
  int rowoff[T];
  bool rowok[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int a = 16 * t + r;
    rowok[t] = a < p.dim(n);
    rowoff[t] = rowok[t] ? a * s_a : 0;
  }
  floatx4 acc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  int eo = (e0 + kq) / n_in, ei = (e0 + kq) - ((e0 + kq) / n_in) * n_in;
  const float* X = p.X;
  auto fetch = [&](int e, float (&xv)[4][T], float (&bv)[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool ok = e + 4 * u + kq < e1;
      bv[u] = ok ? Fo[eo * 16 + r] * Fi[ei * 16 + r] : 0.0f;
      const int off = eo * s_o + ei * s_i;
#pragma unroll
      for (int t = 0; t < T; ++t) xv[u][t] = (ok && rowok[t]) ? X[rowoff[t] + off] : 0.0f;
      ei += 4;
      while (ei >= n_in) {
        ei -= n_in;
        ++eo;
      }
    }
  };
  float xs[D][4][T], bs[D][4];
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (e0 + 16 * d < e1) fetch(e0 + 16 * d, xs[d], bs[d]);
  for (int e = e0; e < e1; e += 16 * D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int ec = e + 16 * d;
      if (ec >= e1) break;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = mfma_f32_16x16x4(xs[d][u][t], bs[d][u], acc[t]);
      if (ec + 16 * D < e1) fetch(ec + 16 * D, xs[d], bs[d]);
    }
  }
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) sm.part[w][t][(4 * kq + j) * 16 + r] = acc[t][j];
  __syncthreads();
  for (int i = threadIdx.x; i < T * 256; i += NW * 64) {
    const int t = i >> 8, k = i & 255;
    float s = sm.part[0][t][k];
    for (int v = 1; v < NW; ++v) s += sm.part[v][t][k];
    sm.part[0][t][k] = s;
  }
  __syncthreads();
}

template <int NW, int D>
__device__ __forceinline__ void mttkrp_depth(int n, int T) {
  if (T <= 1)
    mttkrp_tiles<NW, 1, D>(n);
  else if (T == 2)
    mttkrp_tiles<NW, 2, D>(n);
  else if (T == 3)
    mttkrp_tiles<NW, 3, D>(n);
  else
    mttkrp_tiles<NW, 4, D>(n);
}

template <int NW>
__device__ __noinline__ void mttkrp(int n) {
  const Prob& p = g_sm<NW>.prob;
  const int T = (p.dim(n) + 15) >> 4;
  const int D = g_sm<NW>.kdepth;
  if (D == 1)
    mttkrp_depth<NW, 1>(n, T);
  else if (D == 4)
    mttkrp_depth<NW, 4>(n, T);
  else
    mttkrp_depth<NW, 2>(n, T);
}

// A_n = M V^-1 with V = Hadamard of the other two Grams; V^-1 by Gauss-Jordan
// with partial pivoting (FP64).  Returns 0, or kStSingular.
template <int NW>
__device__ __noinline__ int solve(int n) {
  Smem<NW>& sm = g_sm<NW>;
  const Prob& p = sm.prob;
  const int R = p.R, W = 2 * R;
  const int m1 = (n == 0) ? 1 : 0, m2 = (n == 2) ? 1 : 2;
  for (int i = threadIdx.x; i < R * W; i += NW * 64) {
    const int r = i / W, c = i - r * W;
    sm.aug[r][c] = (c < R) ? sm.G[m1][r * 16 + c] * sm.G[m2][r * 16 + c] : (c - R == r ? 1.0 : 0.0);
  }
  __syncthreads();
  int status = kStOk;
  for (int k = 0; k < R; ++k) {
    int piv = k;
    double best = fabs(sm.aug[k][k]);
    for (int i = k + 1; i < R; ++i) {
      const double v = fabs(sm.aug[i][k]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    if (!(best > 0.0)) {
      status = kStSingular;
      break;
    }
    const double inv = 1.0 / sm.aug[piv][k];
    // new entries for columns k..W-1 of every row (rows k and piv swap)
    // R * (W - k) <= 512 entries: at most 512 / (NW * 64) <= 4 per thread
    constexpr int kPer = (512 + NW * 64 - 1) / (NW * 64);
    double nv[kPer];
    const int span = W - k;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int i = threadIdx.x + j * NW * 64;
      nv[j] = 0.0;
      if (i < R * span) {
        const int row = i / span, c = k + (i - row * span);
        const double pr = sm.aug[piv][c] * inv;
        if (row == k) {
          nv[j] = pr;
        } else {
          const int src = (row == piv) ? k : row;
          nv[j] = sm.aug[src][c] - sm.aug[src][k] * pr;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int i = threadIdx.x + j * NW * 64;
      if (i < R * span) {
        const int row = i / span, c = k + (i - row * span);
        sm.aug[row][c] = nv[j];
      }
    }
    __syncthreads();
  }
  if (status != kStOk) return status;
  const float* M = &sm.part[0][0][0];
  float* F = sm.F[n];
  for (int i = threadIdx.x; i < kBFac; i += NW * 64) {
    const int a = i >> 4, r = i & 15;
    float v = 0.0f;
    if (a < p.dim(n) && r < R) {
      double acc = 0.0;
      for (int s = 0; s < R; ++s) acc += (double)M[a * 16 + s] * sm.aug[s][R + r];
      v = (float)acc;
    }
    F[i] = v;
  }
  __syncthreads();
  gram<NW>(n);
  return kStOk;
}

// ||[[F]]||^2 from the Grams and <X, [[F]]> from M of mode `mode` (in part[0]).
template <int NW>
__device__ double rel_error(int mode, double normx2) {
  Smem<NW>& sm = g_sm<NW>;
  const Prob& p = sm.prob;
  const float* M = &sm.part[0][0][0];
  const float* F = sm.F[mode];
  double ip = 0.0, nm = 0.0;
  // rows >= I of M hold stale values; only the first I_mode rows are read
  for (int i = threadIdx.x; i < p.dim(mode) * 16; i += NW * 64) ip += (double)M[i] * (double)F[i];
  for (int i = threadIdx.x; i < 256; i += NW * 64) nm += sm.G[0][i] * sm.G[1][i] * sm.G[2][i];
  ip = block_sum<NW>(ip);
  nm = block_sum<NW>(nm);
  return sqrt(fabs(normx2 + nm - 2.0 * ip)) / sqrt(normx2);
}

// One ALS sweep over the three modes.  Returns the relative error, or minus
// the status code (singular system, non-finite error).
template <int NW>
__device__ __noinline__ double sweep(double normx2) {
  for (int n = 0; n < 3; ++n) {
    mttkrp<NW>(n);
    const int st = solve<NW>(n);
    if (st != kStOk) return -(double)st;
  }
  const double e = rel_error<NW>(2, normx2);
  if (!(e == e) || e > 1e300) return -(double)kStNonFinite;
  return e;
}

__device__ inline int sweep_status(double e) { return e < 0.0 ? (int)(-e) : kStOk; }

// Error of the current factors by one extra tensor pass (mode-0 MTTKRP): the
// reference line search judges its extrapolated point by a full error
// evaluation (_cp.py:456-458).
template <int NW>
__device__ __noinline__ double full_error(double normx2) {
  grams_all<NW>();
  mttkrp<NW>(0);
  return rel_error<NW>(0, normx2);
}

template <int NW>
__device__ __noinline__ double sqnorm() {
  const Prob& p = g_sm<NW>.prob;
  double a0 = 0.0, a1 = 0.0;
  int64_t i = threadIdx.x;
  for (; i + NW * 64 < p.N; i += 2 * NW * 64) {
    const double x = p.X[i], y = p.X[i + NW * 64];
    a0 += x * x;
    a1 += y * y;
  }
  if (i < p.N) {
    const double x = p.X[i];
    a0 += x * x;
  }
  return block_sum<NW>(a0 + a1);
}

template <int NW>
__device__ void run_problem(const BatchArgs& a, int pid) {
  Smem<NW>& sm = g_sm<NW>;
  const long long t_start = wall_clock64();
  if (threadIdx.x == 0) {
    Prob& p = sm.prob;
    p.X = a.X + a.x_off[pid];
    p.I0 = a.shape[3 * pid];
    p.I1 = a.shape[3 * pid + 1];
    p.I2 = a.shape[3 * pid + 2];
    p.R = a.rank[pid];
    p.N = (int64_t)p.I0 * p.I1 * p.I2;
    p.fac = a.factors + a.f_off[pid];
    p.prev = a.f_prev + a.f_off[pid];
    p.save = a.f_save + a.f_off[pid];
  }
  __syncthreads();
  const Prob& p = sm.prob;
  const int H = a.max_sweeps;
  double* herr = a.hist_err ? a.hist_err + (int64_t)pid * H : nullptr;
  float* halpha = a.hist_alpha ? a.hist_alpha + (int64_t)pid * H : nullptr;
  int* hacc = a.hist_acc ? a.hist_acc + (int64_t)pid * H : nullptr;
  const bool lead = threadIdx.x == 0;

  const double normx2 = sqnorm<NW>();
  load_factors<NW>(p.fac);
  grams_all<NW>();
  int status = kStOk;
  int sweeps = 0, passes = 1, n_extrap = 0, n_accept = 0;
  double err = 1.0;
  auto record = [&](int s, double e, float alpha, int acc) {
    if (lead && herr) {
      herr[s] = e;
      halpha[s] = alpha;
      hacc[s] = acc;
    }
  };

  if (a.policy == kPlain) {
    double prev = 0.0;
    for (int it = 0; it < a.max_sweeps; ++it) {
      const double e = sweep<NW>(normx2);
      if ((status = sweep_status(e))) break;
      ++sweeps;
      passes += 3;
      record(it, e, 0.0f, 1);
      err = e;
      if (it >= 1 && fabs(prev - e) < a.tol) break;
      prev = e;
    }
  } else if (a.policy == kRefLineSearch) {
    double acc_pow = 2.0;
    int acc_fail = 0;
    double r_last = 0.0, r_prev = 0.0;  // rec_errors[-1], rec_errors[-2]
    int n_rec = 0;
    for (int it = 0; it < a.max_sweeps; ++it) {
      if (it % 2 == 0) store_factors<NW>(p.prev);
      double e = sweep<NW>(normx2);
      if ((status = sweep_status(e))) break;
      ++sweeps;
      passes += 3;
      const bool line = (it % 2 == 0) && it > 5;
      float jump_rec = 0.0f;
      int acc_rec = 1;
      if (line) {
        const double jump = pow((double)it, 1.0 / acc_pow);
        store_factors<NW>(p.save);
        extrapolate<NW>(p.prev, (float)(jump - 1.0));
        const double e_new = full_error<NW>(normx2);
        ++passes;
        ++n_extrap;
        jump_rec = (float)jump;
        if (e_new < r_last) {
          e = e_new;
          acc_fail = 0;
          ++n_accept;
          // the Grams of the accepted point are current (full_error built them)
        } else {
          load_factors<NW>(p.save);
          grams_all<NW>();
          acc_rec = 0;
          if (++acc_fail == 4) {
            acc_pow += 1.0;
            acc_fail = 0;
          }
        }
      }
      record(it, e, jump_rec, acc_rec);
      err = e;
      if (!line) {
        r_prev = r_last;
        r_last = e;
        ++n_rec;
        if (it >= 1 && n_rec >= 2 && fabs(r_prev - r_last) < a.tol) break;
      }
    }
  } else {
    // learned / fixed step, or probing
    const bool probe = a.policy == kProbe;
    int s = 0;
    int dec = 0;
    double prev = 0.0;
    bool done = false;
    for (; s < a.max_sweeps && s < a.warmup; ++s) {
      const double e = sweep<NW>(normx2);
      if ((status = sweep_status(e))) break;
      ++sweeps;
      passes += 3;
      record(s, e, 0.0f, 1);
      err = e;
      if (s >= 1 && fabs(prev - e) < a.tol) {
        done = true;
        ++s;
        break;
      }
      prev = e;
    }
    // After a rejected step the next `skip` plain sweeps are not followed by
    // an extrapolation; the pause doubles with consecutive rejections (up to
    // 8) and resets on acceptance.  Not used when probing.
    int backoff = 0, skip = 0;
    while (!done && !status && s < a.max_sweeps) {
      store_factors<NW>(p.prev);
      const double e = sweep<NW>(normx2);
      if ((status = sweep_status(e))) break;
      ++sweeps;
      passes += 3;
      record(s, e, 0.0f, 1);
      err = e;
      if (s >= 1 && fabs(prev - e) < a.tol) break;
      prev = e;
      if (s + 1 >= a.max_sweeps) break;
      if (skip > 0) {
        --skip;
        ++s;
        continue;
      }
      // the features read the history written by thread 0 (same thread)
      if (lead) extrap_features(herr, halpha, hacc, s, sm.feat);
      __syncthreads();
      float alpha = a.alpha_fixed;
      store_factors<NW>(p.save);
      const int d = dec - a.first_dec;
      if (probe && d >= 0 && d < a.max_dec) {
        if (threadIdx.x < kMlpIn) a.feats[((int64_t)pid * a.max_dec + d) * kMlpIn + threadIdx.x] = sm.feat[threadIdx.x];
        int best = 0;
        double best_e = 0.0;
        for (int g = 0; g < a.n_probe; ++g) {
          extrapolate<NW>(p.prev, a.probe_alphas[g]);
          grams_all<NW>();
          double eg = sweep<NW>(normx2);
          passes += 3;
          if (sweep_status(eg)) eg = 1e300;
          if (lead) a.probe_err[((int64_t)pid * a.max_dec + d) * a.n_probe + g] = eg;
          if (g == 0 || eg < best_e) {
            best_e = eg;
            best = g;
          }
          load_factors<NW>(p.save);
        }
        alpha = a.probe_alphas[best];
      } else if (a.mlp) {
        const float y = mlp_alpha_block(a.mlp, sm.feat, sm.h1, sm.hpart, a.alpha_max);
        if (lead) sm.iscal[0] = __float_as_int(y);
        __syncthreads();
        alpha = __int_as_float(sm.iscal[0]);
      }
      ++dec;
      extrapolate<NW>(p.prev, alpha);
      grams_all<NW>();
      const double e2 = sweep<NW>(normx2);
      const int st = sweep_status(e2);
      ++sweeps;
      passes += 3;
      ++n_extrap;
      ++s;
      if (st == kStOk && e2 < e) {
        ++n_accept;
        record(s, e2, alpha, 1);
        err = e2;
        prev = e2;
        backoff = 0;
      } else {
        load_factors<NW>(p.save);
        grams_all<NW>();
        record(s, e, alpha, 0);
        err = e;
        prev = e;
        if (!probe) {
          backoff = backoff ? min(2 * backoff, 8) : 1;
          skip = backoff;
        }
      }
      ++s;
    }
  }
  store_factors<NW>(p.fac);
  if (lead) {
    a.sweeps_out[pid] = sweeps;
    a.passes_out[pid] = passes;
    a.err_out[pid] = err;
    a.n_extrap_out[pid] = n_extrap;
    a.n_accept_out[pid] = n_accept;
    a.status_out[pid] = status;
    if (a.ticks_out) a.ticks_out[pid] = wall_clock64() - t_start;
  }
  __syncthreads();
}

template <int NW>
__global__ void __launch_bounds__(NW * 64) batched_cp_kernel(BatchArgs a) {
  Smem<NW>& sm = g_sm<NW>;
  if (threadIdx.x == 0) sm.kdepth = a.kdepth;
  while (true) {
    if (threadIdx.x == 0) sm.iscal[1] = atomicAdd(a.queue, 1);
    __syncthreads();
    const int slot = sm.iscal[1];
    __syncthreads();
    if (slot >= a.P) break;
    run_problem<NW>(a, a.order[slot]);
  }
}

__global__ void reset_queue_kernel(int* q) {
  if (threadIdx.x == 0) q[0] = 0;
}

}  // namespace
}  // namespace tla

using namespace tla;

extern "C" int tla_batched_cp_max_blocks_per_cu(int waves) {
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
  if (waves == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batched_cp_kernel<2>, 128, 0);
  if (waves == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batched_cp_kernel<4>, 256, 0);
  if (waves == 8) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batched_cp_kernel<8>, 512, 0);
  if (e != hipSuccess) {
    set_error("batched_cp: occupancy query failed for %d waves", waves);
    return -1;
  }
  return n;
}

extern "C" int tla_batched_cp_als(int n_problems, const float* X, const int64_t* x_off, const int* shapes,
                                  const int* ranks, float* factors, const int64_t* f_off, float* f_prev, float* f_save,
                                  const int* order, int* queue, int policy, int max_sweeps, double tol, int warmup,
                                  const float* mlp_params, float alpha_fixed, float alpha_max,
                                  const float* probe_alphas, int n_probe, int max_decisions, int first_decision, double* hist_err,
                                  float* hist_alpha, int* hist_acc, float* features, double* probe_err,
                                  int* sweeps_out, int* passes_out, double* err_out, int* n_extrap_out,
                                  int* n_accept_out, int* status_out, long long* ticks_out, int waves, int grid,
                                  int ktile, void* stream) {
  TLA_REQUIRE(n_problems >= 0, "batched_cp: n_problems = %d", n_problems);
  if (n_problems == 0) return TLA_OK;
  TLA_REQUIRE(X && x_off && shapes && ranks && factors && f_off && f_prev && f_save && order && queue,
              "batched_cp: NULL problem pointer");
  TLA_REQUIRE(sweeps_out && passes_out && err_out && n_extrap_out && n_accept_out && status_out,
              "batched_cp: NULL output pointer");
  TLA_REQUIRE(policy >= kPlain && policy <= kProbe, "batched_cp: policy %d", policy);
  TLA_REQUIRE(max_sweeps >= 1, "batched_cp: max_sweeps = %d", max_sweeps);
  TLA_REQUIRE(waves == 2 || waves == 4 || waves == 8, "batched_cp: waves = %d (2, 4 or 8)", waves);
  TLA_REQUIRE(grid >= 1, "batched_cp: grid = %d", grid);
  TLA_REQUIRE(ktile == 16 || ktile == 32 || ktile == 64, "batched_cp: ktile = %d (16, 32 or 64)", ktile);
  if (policy == kLearned || policy == kProbe) {
    TLA_REQUIRE(hist_err && hist_alpha && hist_acc, "batched_cp: the extrapolating policies need the history arrays");
    TLA_REQUIRE(warmup >= 1, "batched_cp: warmup = %d", warmup);
  }
  if (policy == kProbe) {
    TLA_REQUIRE(probe_alphas && features && probe_err && n_probe >= 1 && n_probe <= kMaxProbe && max_decisions >= 1 &&
                    first_decision >= 0,
                "batched_cp: probing needs 1..%d step values and the feature / probe buffers", kMaxProbe);
  }
  BatchArgs a;
  a.P = n_problems;
  a.X = X;
  a.x_off = x_off;
  a.shape = shapes;
  a.rank = ranks;
  a.factors = factors;
  a.f_off = f_off;
  a.f_prev = f_prev;
  a.f_save = f_save;
  a.order = order;
  a.queue = queue;
  a.policy = policy;
  a.max_sweeps = max_sweeps;
  a.tol = tol;
  a.warmup = warmup;
  a.mlp = mlp_params;
  a.alpha_fixed = alpha_fixed;
  a.alpha_max = alpha_max;
  a.probe_alphas = probe_alphas;
  a.n_probe = n_probe;
  a.max_dec = max_decisions;
  a.first_dec = first_decision;
  a.hist_err = hist_err;
  a.hist_alpha = hist_alpha;
  a.hist_acc = hist_acc;
  a.feats = features;
  a.probe_err = probe_err;
  a.sweeps_out = sweeps_out;
  a.passes_out = passes_out;
  a.err_out = err_out;
  a.n_extrap_out = n_extrap_out;
  a.n_accept_out = n_accept_out;
  a.status_out = status_out;
  a.ticks_out = ticks_out;
  a.kdepth = ktile / 16;
  hipStream_t st = (hipStream_t)stream;
  reset_queue_kernel<<<1, 64, 0, st>>>(queue);
  TLA_LAUNCH_CHECK();
  const unsigned g = (unsigned)grid;
  if (waves == 2) batched_cp_kernel<2><<<g, 128, 0, st>>>(a);
  if (waves == 4) batched_cp_kernel<4><<<g, 256, 0, st>>>(a);
  if (waves == 8) batched_cp_kernel<8><<<g, 512, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
