// Fused MTTKRP for gfx950: out = unfold(X, n) @ khatri_rao(A_{m != n}, w).
//
// The Khatri-Rao product is never formed.  Three kernel families:
//
//  * mttkrp_inner  (output mode is not the contiguous one): the tensor is
//    viewed as X[a, i, b, k] (a = modes before the output mode, i = output
//    mode, b = modes between it and the contiguous mode, k = contiguous
//    mode).  For each (a, b) pair a wave computes D[i, r] = sum_k X[a,i,b,k]
//    F_k[k, r] on MFMA tiles of 16 rows i, with the F_k panel staged in LDS
//    (transposed, so each lane reads 8 consecutive k with one 16/32-byte LDS
//    load), and accumulates s_ab[r] * D[i, r] where s_ab is the Hadamard
//    product of the factor rows of a and b (and the weights).  Each lane
//    reads X with 16/32-byte vector loads; four lanes cover 128 contiguous
//    bytes of one row.
//
//  * mttkrp_last   (output mode is the contiguous one, or a contiguous
//    suffix of modes): out[k, r] = sum_p X[p, k] W[p, r] with the KR row
//    W[p, r] = Gm[p / J, r] * F_J[p % J, r] formed in registers.  X is the
//    MFMA B operand: lane (c, g) loads 4 consecutive columns
//    X[p, k0 + 4c .. 4c+3] (one 16-byte load, 256 contiguous bytes per row
//    per wave), and MFMA number t uses column k0 + 4c + t -- a permutation
//    of the output columns undone when the tile is written.
//
//  * mttkrp_generic  plain vector-ALU fallback for any shape / rank / dtype.
//
// Precisions: FP32 (v_mfma_f32_16x16x4_f32), BF16 (v_mfma_f32_16x16x32_bf16
// on RNE-rounded operands) and BF16X3 (x = hi + lo split of both operands,
// hi*hi + hi*lo + lo*hi, FP32 accumulation).  Partial sums of the split
// reduction are summed in a fixed order by mttkrp_reduce (deterministic).
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>

#include "common.h"

namespace tla {

// ---------------------------------------------------------------------------
// Generic vector-ALU kernel
// ---------------------------------------------------------------------------
struct GenericArgs {
  int ndim, mode, R;
  int64_t shape[kMaxOrder];
  int64_t In, Pb, Q;
  const void* fac[kMaxOrder];
  const void* weights;
};

template <typename T, typename Acc, int RB>
__global__ __launch_bounds__(256) void mttkrp_generic_kernel(const T* __restrict__ X, GenericArgs args,
                                                             T* __restrict__ out) {
  const int64_t i = blockIdx.x;
  const int r0 = blockIdx.y * RB;
  Acc acc[RB];
#pragma unroll
  for (int rr = 0; rr < RB; ++rr) acc[rr] = 0;
  for (int64_t q = threadIdx.x; q < args.Q; q += blockDim.x) {
    const int64_t a = q / args.Pb, b = q - a * args.Pb;
    const Acc x = (Acc)X[(a * args.In + i) * args.Pb + b];
    int64_t idx[kMaxOrder];
    int64_t t = q;
#pragma unroll
    for (int m = kMaxOrder - 1; m >= 0; --m) {
      if (m < args.ndim && m != args.mode) {
        idx[m] = t % args.shape[m];
        t /= args.shape[m];
      }
    }
#pragma unroll
    for (int rr = 0; rr < RB; ++rr) {
      const int r = r0 + rr;
      if (r < args.R) {
        Acc p = x;
#pragma unroll
        for (int m = 0; m < kMaxOrder; ++m)
          if (m < args.ndim && m != args.mode) p *= (Acc)((const T*)args.fac[m])[idx[m] * args.R + r];
        acc[rr] += p;
      }
    }
  }
  __shared__ Acc red[4][RB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int rr = 0; rr < RB; ++rr) {
    const Acc v = wave_sum(acc[rr]);
    if (lane == 0) red[wave][rr] = v;
  }
  __syncthreads();
  if (threadIdx.x < RB) {
    const int r = r0 + threadIdx.x;
    if (r < args.R) {
      Acc s = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
      if (args.weights) s *= (Acc)((const T*)args.weights)[r];
      out[i * args.R + r] = (T)s;
    }
  }
}

// ---------------------------------------------------------------------------
// MFMA inner kernel
// ---------------------------------------------------------------------------
struct InnerArgs {
  const void* X;
  int64_t K, In, Pa, Pb;
  int64_t Sa, Si, Sb;
  int R;
  int na, nb;
  int64_t asz[kMaxOrder], bsz[kMaxOrder];
  const float* afac[kMaxOrder];
  const float* bfac[kMaxOrder];
  const float* weights;
  const float* Fk;
  float* partial;
  int panel, n_panels, ab_splits;
  int64_t pairs_per_wave;
};

template <int RT>
__device__ __forceinline__ void pair_scale(const InnerArgs& args, int64_t a, int64_t b, int c, float s[RT]) {
  const int R = args.R;
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) {
    const int r = nb * 16 + c;
    s[nb] = (r < R) ? (args.weights ? args.weights[r] : 1.0f) : 0.0f;
  }
#pragma unroll
  for (int m = kMaxOrder - 1; m >= 0; --m) {
    if (m < args.na) {
      const int64_t idx = a % args.asz[m];
      a /= args.asz[m];
      const float* f = args.afac[m] + idx * R;
#pragma unroll
      for (int nb = 0; nb < RT; ++nb) {
        const int r = nb * 16 + c;
        if (r < R) s[nb] *= f[r];
      }
    }
  }
#pragma unroll
  for (int m = kMaxOrder - 1; m >= 0; --m) {
    if (m < args.nb) {
      const int64_t idx = b % args.bsz[m];
      b /= args.bsz[m];
      const float* f = args.bfac[m] + idx * R;
#pragma unroll
      for (int nb = 0; nb < RT; ++nb) {
        const int r = nb * 16 + c;
        if (r < R) s[nb] *= f[r];
      }
    }
  }
}

template <typename XT>
__device__ __forceinline__ floatx8 load_x8(const XT* p);
template <>
__device__ __forceinline__ floatx8 load_x8<float>(const float* p) {
  const floatx4 lo = *(const floatx4*)p;
  const floatx4 hi = *(const floatx4*)(p + 4);
  floatx8 v;
  v[0] = lo[0]; v[1] = lo[1]; v[2] = lo[2]; v[3] = lo[3];
  v[4] = hi[0]; v[5] = hi[1]; v[6] = hi[2]; v[7] = hi[3];
  return v;
}

// LDS row pitch of the staged factor panel (elements); the pad keeps the 16
// lanes that read 16 different columns off a single bank.
__host__ __device__ constexpr int inner_pitch_f32(int panel) { return panel + 4; }
__host__ __device__ constexpr int inner_pitch_bf16(int panel) { return panel + 8; }

template <int PREC, typename XT, int RT>
__global__ __launch_bounds__(256) void mttkrp_inner_kernel(InnerArgs args) {
  constexpr int RP = RT * 16;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int panel_id = blockIdx.y % args.n_panels;
  const int split = blockIdx.y / args.n_panels;
  const int64_t kb = (int64_t)panel_id * args.panel;
  const int klen = (int)min((int64_t)args.panel, args.K - kb);
  const int R = args.R;
  const int LDF = inner_pitch_f32(args.panel);
  const int LDB = inner_pitch_bf16(args.panel);
  float* Ct = (float*)smem;
  __bf16* Ch = (__bf16*)smem;
  __bf16* Cl = Ch + RP * LDB;

  // Stage F_k[kb:kb+klen, :] transposed into LDS.
  for (int idx = tid; idx < RP * klen; idx += 256) {
    const int n = idx % RP, kk = idx / RP;
    const float v = (n < R) ? args.Fk[(kb + kk) * R + n] : 0.0f;
    if constexpr (PREC == TLA_PREC_FP32) {
      Ct[n * LDF + kk] = v;
    } else {
      const __bf16 h = (__bf16)v;
      Ch[n * LDB + kk] = h;
      if constexpr (PREC == TLA_PREC_BF16X3) Cl[n * LDB + kk] = (__bf16)(v - (float)h);
    }
  }
  __syncthreads();

  const int64_t i = (int64_t)blockIdx.x * 16 + c;
  const int64_t ri = i < args.In ? i : args.In - 1;
  const int64_t total_pairs = args.Pa * args.Pb;
  const int64_t wave_id = (int64_t)split * 4 + wave;
  const int64_t pbeg = wave_id * args.pairs_per_wave;
  const int64_t pend = min(total_pairs, pbeg + args.pairs_per_wave);
  const XT* X = (const XT*)args.X;

  floatx4 acc[RT];
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int64_t pr = pbeg; pr < pend; ++pr) {
    const int64_t a = pr / args.Pb, b = pr - a * args.Pb;
    float s[RT];
    pair_scale<RT>(args, a, b, c, s);
    const XT* row = X + a * args.Sa + ri * args.Si + b * args.Sb + kb + 8 * g;
    floatx4 d[RT];
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) d[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < klen; kk += 32) {
      if constexpr (PREC == TLA_PREC_FP32) {
        const floatx8 av = load_x8<XT>(row + kk);
#pragma unroll
        for (int nb = 0; nb < RT; ++nb) {
          const floatx8 bv = load_x8<float>(Ct + (nb * 16 + c) * LDF + kk + 8 * g);
#pragma unroll
          for (int v = 0; v < 8; ++v) d[nb] = mfma_f32_16x16x4(av[v], bv[v], d[nb]);
        }
      } else {
        bf16x8 ah, al;
        if constexpr (sizeof(XT) == 2) {
          ah = *(const bf16x8*)(row + kk);
        } else {
          const floatx8 av = load_x8<XT>(row + kk);
          if constexpr (PREC == TLA_PREC_BF16X3) split_bf16(av, ah, al);
          else ah = __builtin_convertvector(av, bf16x8);
        }
#pragma unroll
        for (int nb = 0; nb < RT; ++nb) {
          const int off = (nb * 16 + c) * LDB + kk + 8 * g;
          const bf16x8 bh = *(const bf16x8*)(Ch + off);
          d[nb] = mfma_bf16_16x16x32(ah, bh, d[nb]);
          if constexpr (PREC == TLA_PREC_BF16X3) {
            const bf16x8 bl = *(const bf16x8*)(Cl + off);
            d[nb] = mfma_bf16_16x16x32(ah, bl, d[nb]);
            d[nb] = mfma_bf16_16x16x32(al, bh, d[nb]);
          }
        }
      }
    }
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) acc[nb] += s[nb] * d[nb];
  }

  // Sum the four waves' tiles (fixed order) and write this block's partial.
  __syncthreads();
  float* red = (float*)smem;  // [4][16][RP]
#pragma unroll
  for (int nb = 0; nb < RT; ++nb)
#pragma unroll
    for (int e = 0; e < 4; ++e) red[(wave * 16 + 4 * g + e) * RP + nb * 16 + c] = acc[nb][e];
  __syncthreads();
  for (int idx = tid; idx < 16 * RP; idx += 256) {
    const int row = idx / RP, r = idx % RP;
    const int64_t oi = (int64_t)blockIdx.x * 16 + row;
    if (oi < args.In && r < R) {
      const float v = (red[(0 * 16 + row) * RP + r] + red[(1 * 16 + row) * RP + r]) +
                      (red[(2 * 16 + row) * RP + r] + red[(3 * 16 + row) * RP + r]);
      args.partial[((int64_t)blockIdx.y * args.In + oi) * R + r] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// MFMA last-mode kernel
// ---------------------------------------------------------------------------
struct LastArgs {
  const void* X;
  int64_t K, P, J;
  int R;
  const float* Fj;  // J x R
  const float* Gm;  // (P / J) x R : weights * KR of the other leading modes
  float* partial;   // [gridDim.y][K][R]
  int64_t rows_per_wave;
};

constexpr int kLastWaves = 8;

template <typename XT>
__device__ __forceinline__ floatx4 load_x4(const XT* p);
template <>
__device__ __forceinline__ floatx4 load_x4<float>(const float* p) {
  return *(const floatx4*)p;
}
template <>
__device__ __forceinline__ floatx4 load_x4<__bf16>(const __bf16* p) {
  return __builtin_convertvector(*(const bf16x4*)p, floatx4);
}

template <int PREC, typename XT, int RT>
__global__ __launch_bounds__(512) void mttkrp_last_kernel(LastArgs args) {
  constexpr int RP = RT * 16;
  constexpr int STEP = (PREC == TLA_PREC_FP32) ? 4 : 32;
  __shared__ float red[64 * RP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int R = args.R;
  const int64_t K = args.K, J = args.J;
  const int64_t col = (int64_t)blockIdx.x * 64 + 4 * c;
  const bool colok = col < K;
  const int64_t wave_id = (int64_t)blockIdx.y * kLastWaves + wave;
  const int64_t pb = wave_id * args.rows_per_wave;
  const int64_t pe = min(args.P, pb + args.rows_per_wave);
  const XT* X = (const XT*)args.X;

  floatx4 acc[4][RT];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int mb = 0; mb < RT; ++mb) acc[t][mb] = floatx4{0.f, 0.f, 0.f, 0.f};

  // Track (a, j) = divmod(row, J) of this lane's first row incrementally.
  const int64_t row0 = pb + (STEP == 4 ? g : 8 * g);
  int64_t aa = row0 / J, jj = row0 - aa * J;

  for (int64_t p0 = pb; p0 < pe; p0 += STEP) {
    if constexpr (PREC == TLA_PREC_FP32) {
      const int64_t row = p0 + g;
      const bool ok = row < pe;
      floatx4 xv = floatx4{0.f, 0.f, 0.f, 0.f};
      if (ok && colok) xv = load_x4<XT>(X + row * K + col);
      float wv[RT];
#pragma unroll
      for (int mb = 0; mb < RT; ++mb) {
        const int r = mb * 16 + c;
        wv[mb] = (ok && r < R) ? args.Gm[aa * R + r] * args.Fj[jj * R + r] : 0.0f;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mb = 0; mb < RT; ++mb) acc[t][mb] = mfma_f32_16x16x4(wv[mb], xv[t], acc[t][mb]);
      jj += STEP;
      while (jj >= J) { jj -= J; ++aa; }
    } else {
      // Lane rows p0 + 8g + v, v = 0..7; J % 8 == 0 so they share `aa`.
      floatx4 xv[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int64_t row = p0 + 8 * g + v;
        xv[v] = floatx4{0.f, 0.f, 0.f, 0.f};
        if (row < pe && colok) xv[v] = load_x4<XT>(X + row * K + col);
      }
      const bool okrow = (p0 + 8 * g) < pe;
      bf16x8 wh[RT], wl[RT];
#pragma unroll
      for (int mb = 0; mb < RT; ++mb) {
        const int r = mb * 16 + c;
        floatx8 w8;
        const float gm = (okrow && r < R) ? args.Gm[aa * R + r] : 0.0f;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          const bool ok = (p0 + 8 * g + v) < pe && r < R;
          w8[v] = ok ? gm * args.Fj[(jj + v) * R + r] : 0.0f;
        }
        if constexpr (PREC == TLA_PREC_BF16X3) split_bf16(w8, wh[mb], wl[mb]);
        else wh[mb] = __builtin_convertvector(w8, bf16x8);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        floatx8 x8;
#pragma unroll
        for (int v = 0; v < 8; ++v) x8[v] = xv[v][t];
        bf16x8 xh, xl;
        if constexpr (PREC == TLA_PREC_BF16X3) split_bf16(x8, xh, xl);
        else xh = __builtin_convertvector(x8, bf16x8);
#pragma unroll
        for (int mb = 0; mb < RT; ++mb) {
          acc[t][mb] = mfma_bf16_16x16x32(wh[mb], xh, acc[t][mb]);
          if constexpr (PREC == TLA_PREC_BF16X3) {
            acc[t][mb] = mfma_bf16_16x16x32(wh[mb], xl, acc[t][mb]);
            acc[t][mb] = mfma_bf16_16x16x32(wl[mb], xh, acc[t][mb]);
          }
        }
      }
      jj += STEP;
      while (jj >= J) { jj -= J; ++aa; }
    }
  }

  // acc[t][mb][e] = out[k0 + 4c + t][mb*16 + 4g + e]; sum the waves in order.
  for (int w = 0; w < kLastWaves; ++w) {
    if (wave == w) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mb = 0; mb < RT; ++mb)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int idx = (4 * c + t) * RP + mb * 16 + 4 * g + e;
            red[idx] = (w == 0 ? 0.0f : red[idx]) + acc[t][mb][e];
          }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < 64 * RP; idx += 512) {
    const int kc = idx / RP, r = idx % RP;
    const int64_t k = (int64_t)blockIdx.x * 64 + kc;
    if (k < K && r < R) args.partial[((int64_t)blockIdx.y * K + k) * R + r] = red[idx];
  }
}

// ---------------------------------------------------------------------------
// Split reduction and helpers
// ---------------------------------------------------------------------------
__global__ void mttkrp_reduce_kernel(const float* __restrict__ partial, int64_t n, int splits,
                                     float* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += partial[(int64_t)k * n + e];
  out[e] = s;
}

// Gm[a, r] = w[r] * prod_m A_m[a_m, r] over up to kMaxOrder leading modes
// (a = row-major index over those modes); a single row of weights when the
// list is empty.
struct GmArgs {
  int n;
  int64_t sz[kMaxOrder];
  const float* fac[kMaxOrder];
  const float* weights;
  int R;
  int64_t rows;
};
__global__ void build_gm_kernel(GmArgs args, float* __restrict__ gm) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= args.rows * args.R) return;
  const int r = (int)(e % args.R);
  int64_t a = e / args.R;
  float v = args.weights ? args.weights[r] : 1.0f;
#pragma unroll
  for (int m = kMaxOrder - 1; m >= 0; --m) {
    if (m < args.n) {
      const int64_t idx = a % args.sz[m];
      a /= args.sz[m];
      v *= args.fac[m][idx * args.R + r];
    }
  }
  gm[e] = v;
}

// ---------------------------------------------------------------------------
// Host planning
// ---------------------------------------------------------------------------
enum class Path { Generic, Inner, Last };

struct Plan {
  Path path;
  int RT;
  // inner
  InnerArgs in;
  dim3 grid;
  size_t smem;
  // last
  LastArgs last;
  GmArgs gm;
  int64_t gm_rows;
  // common
  int64_t out_rows;
  int splits;
  size_t partial_elems;
  size_t ws_bytes;
};

static bool prec_ok(int x_dtype, int precision) {
  if (x_dtype == TLA_BF16) return precision == TLA_PREC_BF16;
  if (x_dtype == TLA_F32)
    return precision == TLA_PREC_FP32 || precision == TLA_PREC_BF16 || precision == TLA_PREC_BF16X3;
  return false;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Plan the contraction that keeps the contiguous block of modes
// [kb, ke) and contracts every other mode with its factor.  Standard MTTKRP
// of mode n is kb = n, ke = n + 1.
static int make_plan(int x_dtype, int ndim, const int64_t* shape, int R, int kb, int ke,
                     int precision, Plan* p) {
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "tensor order %d not in [2, %d]", ndim, kMaxOrder);
  TLA_REQUIRE(R >= 1, "rank must be >= 1");
  TLA_REQUIRE(0 <= kb && kb < ke && ke <= ndim, "bad kept-mode range [%d, %d)", kb, ke);
  for (int m = 0; m < ndim; ++m) TLA_REQUIRE(shape[m] >= 1, "empty mode %d", m);
  *p = Plan{};
  const bool standard = (ke == kb + 1);
  const bool mfma = prec_ok(x_dtype, precision) && R <= 64;
  const int RT = (R + 15) / 16;
  p->RT = RT;
  const int64_t K = shape[ndim - 1];
  if (ke < ndim) {
    // Output block is not the contiguous mode: inner kernel.
    TLA_REQUIRE(standard || kb == 0, "partial contraction must keep a prefix or a suffix");
    int64_t In = 1;
    for (int m = kb; m < ke; ++m) In *= shape[m];
    int64_t Pa = 1, Pb = 1;
    for (int m = 0; m < kb; ++m) Pa *= shape[m];
    for (int m = ke; m < ndim - 1; ++m) Pb *= shape[m];
    p->out_rows = In;
    if (!(mfma && K % 32 == 0)) {
      TLA_REQUIRE(standard, "partial contraction needs an MFMA-eligible shape");
      TLA_REQUIRE(x_dtype != TLA_BF16, "bf16 tensor copy needs K %% 32 == 0 and R <= 64");
      p->path = Path::Generic;
      p->ws_bytes = 0;
      return TLA_OK;
    }
    p->path = Path::Inner;
    InnerArgs& a = p->in;
    a.K = K; a.In = In; a.Pa = Pa; a.Pb = Pb;
    a.Sb = K; a.Si = Pb * K; a.Sa = In * Pb * K;
    a.R = R;
    a.na = kb; a.nb = ndim - 1 - ke;
    for (int m = 0; m < kb; ++m) a.asz[m] = shape[m];
    for (int m = ke; m < ndim - 1; ++m) a.bsz[m - ke] = shape[m];
    a.panel = (int)std::min<int64_t>(K, 128);
    a.n_panels = ceil_div(K, a.panel);
    const int itiles = ceil_div(In, 16);
    const int64_t pairs = Pa * Pb;
    int64_t ab = std::max<int64_t>(1, (2048 + (int64_t)itiles * a.n_panels - 1) / ((int64_t)itiles * a.n_panels));
    ab = std::min<int64_t>(ab, (pairs + 3) / 4);
    ab = std::min<int64_t>(ab, 65535 / a.n_panels);
    a.ab_splits = (int)ab;
    a.pairs_per_wave = (pairs + 4 * ab - 1) / (4 * ab);
    p->grid = dim3(itiles, a.n_panels * a.ab_splits, 1);
    p->splits = a.n_panels * a.ab_splits;
    const size_t stage = (precision == TLA_PREC_FP32)
                             ? size_t(RT * 16) * inner_pitch_f32(a.panel) * 4
                             : size_t(2) * RT * 16 * inner_pitch_bf16(a.panel) * 2;
    p->smem = std::max(stage, size_t(4 * 16 * RT * 16 * 4));
    p->partial_elems = size_t(p->splits) * In * R;
    p->ws_bytes = align256(p->partial_elems * 4);
    return TLA_OK;
  }
  // Output block is a suffix including the contiguous mode: last kernel.
  int64_t Kc = 1;
  for (int m = kb; m < ndim; ++m) Kc *= shape[m];
  int64_t P = 1;
  for (int m = 0; m < kb; ++m) P *= shape[m];
  p->out_rows = Kc;
  const int64_t J = kb >= 1 ? shape[kb - 1] : 1;
  const bool shape_ok = Kc % 4 == 0 && kb >= 1 &&
                        (precision == TLA_PREC_FP32 || J % 8 == 0);
  if (!(mfma && shape_ok)) {
    TLA_REQUIRE(standard, "partial contraction needs an MFMA-eligible shape");
    TLA_REQUIRE(x_dtype != TLA_BF16, "bf16 tensor copy needs an MFMA-eligible shape");
    p->path = Path::Generic;
    p->ws_bytes = 0;
    return TLA_OK;
  }
  p->path = Path::Last;
  LastArgs& l = p->last;
  l.K = Kc; l.P = P; l.J = J; l.R = R;
  GmArgs& gm = p->gm;
  gm.n = kb - 1;
  gm.R = R;
  int64_t gm_rows = 1;
  for (int m = 0; m < kb - 1; ++m) { gm.sz[m] = shape[m]; gm_rows *= shape[m]; }
  gm.rows = gm_rows;
  p->gm_rows = gm_rows;
  const int ktiles = ceil_div(Kc, 64);
  const int step = precision == TLA_PREC_FP32 ? 4 : 32;
  int64_t S = std::max<int64_t>(1, (1024 + ktiles - 1) / ktiles);
  const int64_t max_s = std::max<int64_t>(1, P / (kLastWaves * step));
  S = std::min<int64_t>(S, max_s);
  S = std::min<int64_t>(S, 65535);
  int64_t rpw = (P + S * kLastWaves - 1) / (S * kLastWaves);
  rpw = (rpw + step - 1) / step * step;
  S = (P + rpw * kLastWaves - 1) / (rpw * kLastWaves);
  l.rows_per_wave = rpw;
  p->grid = dim3(ktiles, (unsigned)S, 1);
  p->splits = (int)S;
  p->partial_elems = size_t(S) * Kc * R;
  p->ws_bytes = align256(size_t(gm_rows) * R * 4) + align256(p->partial_elems * 4);
  return TLA_OK;
}

template <int PREC, typename XT>
static int launch_inner(const Plan& p, hipStream_t st) {
  InnerArgs a = p.in;
  switch (p.RT) {
#define TLA_INNER_CASE(RTV)                                                                          \
  case RTV:                                                                                          \
    mttkrp_inner_kernel<PREC, XT, RTV><<<p.grid, 256, p.smem, st>>>(a);                              \
    break;
    TLA_INNER_CASE(1)
    TLA_INNER_CASE(2)
    TLA_INNER_CASE(3)
    TLA_INNER_CASE(4)
#undef TLA_INNER_CASE
    default:
      set_error("rank tile %d unsupported", p.RT);
      return TLA_ERR_UNSUPPORTED;
  }
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

template <int PREC, typename XT>
static int launch_last(const Plan& p, hipStream_t st) {
  LastArgs a = p.last;
  switch (p.RT) {
#define TLA_LAST_CASE(RTV)                                                                           \
  case RTV:                                                                                          \
    mttkrp_last_kernel<PREC, XT, RTV><<<p.grid, 512, 0, st>>>(a);                                    \
    break;
    TLA_LAST_CASE(1)
    TLA_LAST_CASE(2)
    TLA_LAST_CASE(3)
    TLA_LAST_CASE(4)
#undef TLA_LAST_CASE
    default:
      set_error("rank tile %d unsupported", p.RT);
      return TLA_ERR_UNSUPPORTED;
  }
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

template <typename T, typename Acc>
static int launch_generic(const T* X, int ndim, const int64_t* shape, const void* const* factors,
                          const void* weights, int R, int mode, T* out, hipStream_t st) {
  GenericArgs g{};
  g.ndim = ndim; g.mode = mode; g.R = R;
  g.In = shape[mode]; g.Pb = 1; g.Q = 1;
  for (int m = 0; m < ndim; ++m) {
    g.shape[m] = shape[m];
    g.fac[m] = factors[m];
    if (m != mode) g.Q *= shape[m];
    if (m > mode) g.Pb *= shape[m];
  }
  g.weights = weights;
  constexpr int RB = 16;
  dim3 grid((unsigned)g.In, (unsigned)ceil_div(R, RB), 1);
  TLA_REQUIRE(g.In <= 0x7fffffff, "mode too large for the generic path");
  mttkrp_generic_kernel<T, Acc, RB><<<grid, 256, 0, st>>>(X, g, out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

static int run_plan(const Plan& p0, int x_dtype, const void* X, int ndim, const int64_t* shape,
                    const void* const* factors, const void* weights, int R, int kb, int ke,
                    int precision, void* out, void* ws, size_t ws_bytes, hipStream_t st) {
  Plan p = p0;
  if (p.path == Path::Generic) {
    if (x_dtype == TLA_F64)
      return launch_generic<double, double>((const double*)X, ndim, shape, factors, weights, R, kb,
                                            (double*)out, st);
    return launch_generic<float, float>((const float*)X, ndim, shape, factors, weights, R, kb,
                                        (float*)out, st);
  }
  TLA_REQUIRE(ws_bytes >= p.ws_bytes, "workspace too small: %zu < %zu", ws_bytes, p.ws_bytes);
  TLA_REQUIRE(ws != nullptr || p.ws_bytes == 0, "workspace is NULL");
  for (int m = 0; m < ndim; ++m)
    TLA_REQUIRE((m >= kb && m < ke) || factors[m] != nullptr, "factor %d is NULL", m);
  int rc = TLA_OK;
  float* partial;
  if (p.path == Path::Inner) {
    InnerArgs& a = p.in;
    a.X = X;
    for (int m = 0; m < a.na; ++m) a.afac[m] = (const float*)factors[m];
    for (int m = 0; m < a.nb; ++m) a.bfac[m] = (const float*)factors[ke + m];
    a.weights = (const float*)weights;
    a.Fk = (const float*)factors[ndim - 1];
    a.partial = partial = (float*)ws;
    if (x_dtype == TLA_BF16) rc = launch_inner<TLA_PREC_BF16, __bf16>(p, st);
    else if (precision == TLA_PREC_FP32) rc = launch_inner<TLA_PREC_FP32, float>(p, st);
    else if (precision == TLA_PREC_BF16) rc = launch_inner<TLA_PREC_BF16, float>(p, st);
    else rc = launch_inner<TLA_PREC_BF16X3, float>(p, st);
  } else {
    LastArgs& l = p.last;
    float* gm = (float*)ws;
    partial = (float*)((char*)ws + align256(size_t(p.gm_rows) * R * 4));
    GmArgs g = p.gm;
    for (int m = 0; m < g.n; ++m) g.fac[m] = (const float*)factors[m];
    g.weights = (const float*)weights;
    const int64_t ne = p.gm_rows * R;
    build_gm_kernel<<<ceil_div(ne, 256), 256, 0, st>>>(g, gm);
    TLA_LAUNCH_CHECK();
    l.X = X;
    l.Fj = (const float*)factors[kb - 1];
    l.Gm = gm;
    l.partial = partial;
    if (x_dtype == TLA_BF16) rc = launch_last<TLA_PREC_BF16, __bf16>(p, st);
    else if (precision == TLA_PREC_FP32) rc = launch_last<TLA_PREC_FP32, float>(p, st);
    else if (precision == TLA_PREC_BF16) rc = launch_last<TLA_PREC_BF16, float>(p, st);
    else rc = launch_last<TLA_PREC_BF16X3, float>(p, st);
  }
  if (rc != TLA_OK) return rc;
  const int64_t n = p.out_rows * R;
  mttkrp_reduce_kernel<<<ceil_div(n, 256), 256, 0, st>>>(partial, n, p.splits, (float*)out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

}  // namespace tla

using namespace tla;

extern "C" size_t tla_mttkrp_workspace_size(int x_dtype, int ndim, const int64_t* shape, int rank,
                                            int mode, int precision) {
  if (mode < 0 || mode >= ndim) return 0;
  Plan p;
  if (make_plan(x_dtype, ndim, shape, rank, mode, mode + 1, precision, &p) != TLA_OK) return 0;
  return p.ws_bytes;
}

extern "C" int tla_mttkrp(int x_dtype, const void* X, int ndim, const int64_t* shape,
                          const void* const* factors, const void* weights, int rank, int mode,
                          int precision, void* out, void* workspace, size_t workspace_bytes,
                          void* stream) {
  TLA_REQUIRE(x_dtype == TLA_F32 || x_dtype == TLA_F64 || x_dtype == TLA_BF16, "bad dtype %d", x_dtype);
  TLA_REQUIRE(mode >= 0 && mode < ndim, "mode %d out of range for order %d", mode, ndim);
  TLA_REQUIRE(X && out && factors, "NULL pointer argument");
  if (x_dtype == TLA_F64 && precision != TLA_PREC_VALU) precision = TLA_PREC_VALU;
  Plan p;
  int rc = make_plan(x_dtype, ndim, shape, rank, mode, mode + 1, precision, &p);
  if (rc != TLA_OK) return rc;
  return run_plan(p, x_dtype, X, ndim, shape, factors, weights, rank, mode, mode + 1, precision,
                  out, workspace, workspace_bytes, (hipStream_t)stream);
}

extern "C" size_t tla_mttkrp_partial_workspace_size(int x_dtype, int ndim, const int64_t* shape,
                                                    int rank, int keep_begin, int keep_end,
                                                    int precision) {
  Plan p;
  if (make_plan(x_dtype, ndim, shape, rank, keep_begin, keep_end, precision, &p) != TLA_OK) return 0;
  return p.ws_bytes;
}

extern "C" int tla_mttkrp_partial(int x_dtype, const void* X, int ndim, const int64_t* shape,
                                  const void* const* factors, const void* weights, int rank,
                                  int keep_begin, int keep_end, int precision, void* out,
                                  void* workspace, size_t workspace_bytes, void* stream) {
  TLA_REQUIRE(x_dtype == TLA_F32 || x_dtype == TLA_BF16, "partial MTTKRP supports f32 / bf16 tensors");
  TLA_REQUIRE(keep_begin == 0 || keep_end == ndim, "kept modes must be a prefix or a suffix");
  TLA_REQUIRE(!(keep_begin == 0 && keep_end == ndim), "at least one mode must be contracted");
  TLA_REQUIRE(X && out && factors, "NULL pointer argument");
  Plan p;
  int rc = make_plan(x_dtype, ndim, shape, rank, keep_begin, keep_end, precision, &p);
  if (rc != TLA_OK) return rc;
  TLA_REQUIRE(p.path != Path::Generic, "partial MTTKRP needs an MFMA-eligible shape");
  return run_plan(p, x_dtype, X, ndim, shape, factors, weights, rank, keep_begin, keep_end,
                  precision, out, workspace, workspace_bytes, (hipStream_t)stream);
}
