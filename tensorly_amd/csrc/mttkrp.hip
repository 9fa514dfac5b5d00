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
// Precisions: FP32 (v_mfma_f32_16x16x4_f32), BF16 (v_mfma_f32_16x16x32_bf16:
// the tensor RNE-rounded to bf16, or a bf16 copy of it, times the factor
// side split as hi + lo, two products: exact for the rounded tensor up to
// FP32 accumulation, so ALS stays consistent with its own Grams) and BF16X3
// (x = hi + lo split of both operands, hi*hi + hi*lo + lo*hi, FP32
// accumulation).  Partial sums of the split
// reduction are summed in a fixed order by mttkrp_reduce (deterministic).
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <type_traits>

#include "common.h"
#include "inverse.h"

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
      Cl[n * LDB + kk] = (__bf16)(v - (float)h);
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
          const bf16x8 bl = *(const bf16x8*)(Cl + off);
          d[nb] = mfma_bf16_16x16x32(ah, bh, d[nb]);
          d[nb] = mfma_bf16_16x16x32(ah, bl, d[nb]);
          if constexpr (PREC == TLA_PREC_BF16X3) d[nb] = mfma_bf16_16x16x32(al, bh, d[nb]);
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
// MFMA row-tile kernel.  Rows p of X (all modes but the contiguous one,
// flattened) are split into units of U = 16 * tpu consecutive rows; wave u
// computes D = X[rows] @ F_k for its tpu tiles (the F_k panel in LDS is read
// once per k-step and feeds all tpu tiles) and then either
//   EPI_REDUCE: out_partial[u, r] = s_q[r] * sum_{rows} F_j[row % J, r] D[row, r]
//               (q = row / J is constant over a unit because J % U == 0), or
//   EPI_ROWS:   out[row, r] = w[r] * D[row, r]   (dimension-tree contraction).
// ---------------------------------------------------------------------------
constexpr int EPI_REDUCE = 0;
constexpr int EPI_ROWS = 1;
constexpr int kTilesPerWave = 4;

struct TilesArgs {
  const void* X;
  int64_t K, P, J;
  int R;
  int tpu;             // tiles (16 rows) per unit, divides J / 16
  int64_t n_units;
  int panel;
  // s_q = w * prod_{a modes} A_m[q_m] * prod_{b modes} A_m[q_m]; q = (a, i, b)
  int na, nb;
  int64_t In, Pb;      // q = (a * In + i) * Pb + b
  int64_t asz[kMaxOrder], bsz[kMaxOrder];
  const float* afac[kMaxOrder];
  const float* bfac[kMaxOrder];
  const float* weights;
  const float* Fj;     // J x R (EPI_REDUCE)
  const float* Fk;     // K x R
  float* out;          // EPI_REDUCE: [n_units][R]; EPI_ROWS: [P][R]
};

template <int PREC, typename XT, int RT, int EPI>
__global__ __launch_bounds__(256) void mttkrp_tiles_kernel(TilesArgs args) {
  constexpr int RP = RT * 16;
  constexpr int TPW = kTilesPerWave;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int R = args.R;
  const int tpu = args.tpu;
  const int U = 16 * tpu;
  const int64_t unit = (int64_t)blockIdx.x * 4 + wave;
  const bool active = unit < args.n_units;
  const int64_t row0 = unit * U;
  const int LDF = inner_pitch_f32(args.panel);
  const int LDB = inner_pitch_bf16(args.panel);
  float* Ct = (float*)smem;
  __bf16* Ch = (__bf16*)smem;
  __bf16* Cl = Ch + RP * LDB;
  const XT* X = (const XT*)args.X;

  floatx4 d[TPW][RT];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) d[t][nb] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int64_t kb = 0; kb < args.K; kb += args.panel) {
    const int klen = (int)min((int64_t)args.panel, args.K - kb);
    __syncthreads();
    for (int idx = tid; idx < RP * klen; idx += 256) {
      const int n = idx % RP, kk = idx / RP;
      const float v = (n < R) ? args.Fk[(kb + kk) * R + n] : 0.0f;
      if constexpr (PREC == TLA_PREC_FP32) {
        Ct[n * LDF + kk] = v;
      } else {
        const __bf16 h = (__bf16)v;
        Ch[n * LDB + kk] = h;
        Cl[n * LDB + kk] = (__bf16)(v - (float)h);
      }
    }
    __syncthreads();
    if (active) {
      for (int kk = 0; kk < klen; kk += 32) {
        floatx8 av[TPW];
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          if (t < tpu) {
            const XT* p = X + (row0 + 16 * t + c) * args.K + kb + kk + 8 * g;
            if constexpr (sizeof(XT) == 2) {
              av[t] = __builtin_convertvector(*(const bf16x8*)p, floatx8);
            } else {
              av[t] = load_x8<XT>(p);
            }
          }
        }
        if constexpr (PREC == TLA_PREC_FP32) {
#pragma unroll
          for (int nb = 0; nb < RT; ++nb) {
            const floatx8 bv = load_x8<float>(Ct + (nb * 16 + c) * LDF + kk + 8 * g);
#pragma unroll
            for (int t = 0; t < TPW; ++t)
              if (t < tpu)
#pragma unroll
                for (int v = 0; v < 8; ++v) d[t][nb] = mfma_f32_16x16x4(av[t][v], bv[v], d[t][nb]);
          }
        } else {
          bf16x8 ah[TPW], al[TPW];
#pragma unroll
          for (int t = 0; t < TPW; ++t) {
            if (t < tpu) {
              if constexpr (PREC == TLA_PREC_BF16X3) split_bf16(av[t], ah[t], al[t]);
              else ah[t] = __builtin_convertvector(av[t], bf16x8);
            }
          }
#pragma unroll
          for (int nb = 0; nb < RT; ++nb) {
            const int off = (nb * 16 + c) * LDB + kk + 8 * g;
            const bf16x8 bh = *(const bf16x8*)(Ch + off);
            const bf16x8 bl = *(const bf16x8*)(Cl + off);
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
              if (t < tpu) {
                d[t][nb] = mfma_bf16_16x16x32(ah[t], bh, d[t][nb]);
                d[t][nb] = mfma_bf16_16x16x32(ah[t], bl, d[t][nb]);
                if constexpr (PREC == TLA_PREC_BF16X3) d[t][nb] = mfma_bf16_16x16x32(al[t], bh, d[t][nb]);
              }
            }
          }
        }
      }
    }
  }
  if (!active) return;
  // d[t][nb][e] = D[row0 + 16 t + 4 g + e][nb * 16 + c]
  if constexpr (EPI == EPI_ROWS) {
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const int r = nb * 16 + c;
      if (r >= R) continue;
      const float w = args.weights ? args.weights[r] : 1.0f;
#pragma unroll
      for (int t = 0; t < TPW; ++t)
        if (t < tpu)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t row = row0 + 16 * t + 4 * g + e;
            if (row < args.P) args.out[row * R + r] = w * d[t][nb][e];
          }
    }
  } else {
    const int64_t q = row0 / args.J;
    const int64_t j0 = row0 - q * args.J;
    float s[RT];
    {
      int64_t b = q % args.Pb;
      int64_t a = (q / args.Pb) / args.In;
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
#pragma unroll
          for (int nb = 0; nb < RT; ++nb) {
            const int r = nb * 16 + c;
            if (r < R) s[nb] *= args.afac[m][idx * R + r];
          }
        }
      }
#pragma unroll
      for (int m = kMaxOrder - 1; m >= 0; --m) {
        if (m < args.nb) {
          const int64_t idx = b % args.bsz[m];
          b /= args.bsz[m];
#pragma unroll
          for (int nb = 0; nb < RT; ++nb) {
            const int r = nb * 16 + c;
            if (r < R) s[nb] *= args.bfac[m][idx * R + r];
          }
        }
      }
    }
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const int r = nb * 16 + c;
      float acc = 0.0f;
      if (r < R) {
#pragma unroll
        for (int t = 0; t < TPW; ++t)
          if (t < tpu)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc = fmaf(args.Fj[(j0 + 16 * t + 4 * g + e) * R + r], d[t][nb][e], acc);
      }
      acc += __shfl_xor(acc, 16, 64);
      acc += __shfl_xor(acc, 32, 64);
      if (g == 0 && r < R) args.out[unit * R + r] = s[nb] * acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Persistent interleaved row-tile kernel ("stream").  The whole F_k (K x R)
// is staged once per block in LDS; wave w of W visits 16-row tiles
// w, w + W, w + 2W, ... so that at any time the grid reads one narrow window
// of the tensor (measured: ~6.3 TB/s for this pattern versus ~5.4 TB/s when
// each wave owns a contiguous slab).  Loads are double-buffered in
// registers, 64 columns (two MFMA k-steps) per buffer.  Epilogues:
//   SE_ROWS   out[row, r] = w[r] * D[row, r]
//   SE_REDQ   partial[tile, r] = s_q[r] * sum_rows F_j[row % J, r] D[row, r]
//   SE_ACCJ   (output mode = mode N-2, rows of a tile are output rows):
//             acc[rows] += s_q * D, W % (J/16) == 0 keeps the wave on one
//             16-row output block; partial[wave, 16, r].
//   SE_REDQG  SE_REDQ with the J/16 tiles of one q visited by one wave in a
//             row (wave w takes q = w, w + W, ...): the row sums accumulate
//             in registers over the q and one partial per q is stored;
//             F_j is staged in LDS (transposed) instead of riding in the
//             prefetch slot.  Chosen when there are at least 2 W values of q.
// ---------------------------------------------------------------------------
constexpr int SE_ROWS = 0;
constexpr int SE_REDQ = 1;
constexpr int SE_ACCJ = 2;
constexpr int SE_REDQG = 3;

struct StreamArgs {
  const void* X;
  int64_t K, P, J;
  int R;
  int64_t ntiles;
  // s_q = w * prod_{a modes} A_m[q_m] * prod_{b modes} A_m[q_m]; q = (a, i, b)
  int na, nb;
  int64_t In, Pb;
  int64_t asz[kMaxOrder], bsz[kMaxOrder];
  const float* afac[kMaxOrder];
  const float* bfac[kMaxOrder];
  const float* weights;
  const float* Fj;
  const float* Fk;
  const float* Sq;  // (ntiles / (J / 16)) x R block scales, precomputed
  float* out;
  InvJob inv;  // optional side job for the grid's last workgroup
};

template <int RT>
__device__ __forceinline__ void scale_q(const StreamArgs& a, int64_t q, int c, float s[RT]) {
  const int R = a.R;
  int64_t b = q % a.Pb;
  int64_t aa = (q / a.Pb) / a.In;
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) {
    const int r = nb * 16 + c;
    s[nb] = (r < R) ? (a.weights ? a.weights[r] : 1.0f) : 0.0f;
  }
#pragma unroll
  for (int m = kMaxOrder - 1; m >= 0; --m) {
    if (m < a.na) {
      const int64_t idx = aa % a.asz[m];
      aa /= a.asz[m];
#pragma unroll
      for (int nb = 0; nb < RT; ++nb) {
        const int r = nb * 16 + c;
        if (r < R) s[nb] *= a.afac[m][idx * R + r];
      }
    }
  }
#pragma unroll
  for (int m = kMaxOrder - 1; m >= 0; --m) {
    if (m < a.nb) {
      const int64_t idx = b % a.bsz[m];
      b /= a.bsz[m];
#pragma unroll
      for (int nb = 0; nb < RT; ++nb) {
        const int r = nb * 16 + c;
        if (r < R) s[nb] *= a.bfac[m][idx * R + r];
      }
    }
  }
}

// The prefetch slots keep 8 tensor values as loaded (float, or bf16 for a
// bf16 copy): a conversion at load time would make the wave wait for the
// load right away and defeat the prefetch.
template <typename XT> struct StreamVec { using type = floatx8; };
template <> struct StreamVec<__bf16> { using type = bf16x8; };

// 8 consecutive tensor values at p (16-byte aligned); NT: non-temporal
// (streaming) loads, which measured faster for the kernels whose output is
// small, and slower for the T01 partial that writes 33.5 MB while it reads.
template <typename XT, bool NT>
__device__ __forceinline__ typename StreamVec<XT>::type load_stream_x8(const XT* p) {
  if constexpr (sizeof(XT) == 2) {
    return NT ? __builtin_nontemporal_load((const bf16x8*)p) : *(const bf16x8*)p;
  } else {
    const floatx4 lo = NT ? __builtin_nontemporal_load((const floatx4*)p) : *(const floatx4*)p;
    const floatx4 hi = NT ? __builtin_nontemporal_load((const floatx4*)(p + 4)) : *(const floatx4*)(p + 4);
    floatx8 v;
    v[0] = lo[0]; v[1] = lo[1]; v[2] = lo[2]; v[3] = lo[3];
    v[4] = hi[0]; v[5] = hi[1]; v[6] = hi[2]; v[7] = hi[3];
    return v;
  }
}

template <typename XT, bool NT = false>
__device__ __forceinline__ void load_pair(const XT* p, typename StreamVec<XT>::type& v0,
                                          typename StreamVec<XT>::type& v1) {
  v0 = load_stream_x8<XT, NT>(p);
  v1 = load_stream_x8<XT, NT>(p + 32);
}

// A bf16 copy (BF16 only): the loaded values are the A operand as they are.
template <int PREC, int RT>
__device__ __forceinline__ void stream_step(const bf16x8& ah, const float* Ct, const __bf16* Ch, const __bf16* Cl,
                                            int LDF, int LDB, int kk, int c, int g, floatx4 d[RT]) {
  static_assert(PREC == TLA_PREC_BF16, "a bf16 tensor copy runs BF16");
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) {
    const int off = (nb * 16 + c) * LDB + kk + 8 * g;
    const bf16x8 bh = *(const bf16x8*)(Ch + off);
    const bf16x8 bl = *(const bf16x8*)(Cl + off);
    d[nb] = mfma_bf16_16x16x32(ah, bh, d[nb]);
    d[nb] = mfma_bf16_16x16x32(ah, bl, d[nb]);
  }
}

template <int PREC, int RT>
__device__ __forceinline__ void stream_step(const floatx8& av, const float* Ct, const __bf16* Ch, const __bf16* Cl,
                                            int LDF, int LDB, int kk, int c, int g, floatx4 d[RT]) {
  if constexpr (PREC == TLA_PREC_FP32) {
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const floatx8 bv = load_x8<float>(Ct + (nb * 16 + c) * LDF + kk + 8 * g);
#pragma unroll
      for (int v = 0; v < 8; ++v) d[nb] = mfma_f32_16x16x4(av[v], bv[v], d[nb]);
    }
  } else {
    bf16x8 ah, al;
    if constexpr (PREC == TLA_PREC_BF16X3) split_bf16(av, ah, al);
    else ah = __builtin_convertvector(av, bf16x8);
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const int off = (nb * 16 + c) * LDB + kk + 8 * g;
      const bf16x8 bh = *(const bf16x8*)(Ch + off);
      const bf16x8 bl = *(const bf16x8*)(Cl + off);
      d[nb] = mfma_bf16_16x16x32(ah, bh, d[nb]);
      d[nb] = mfma_bf16_16x16x32(ah, bl, d[nb]);
      if constexpr (PREC == TLA_PREC_BF16X3) d[nb] = mfma_bf16_16x16x32(al, bh, d[nb]);
    }
  }
}

constexpr int kStreamWaves = 8;  // 512-thread blocks share one LDS copy of F_k

// One ring slot: a 64-column chunk of a 16-row tile plus, when the chunk is
// the tile's last, the tile's epilogue operands (loaded in the same batch so
// the in-order return of vector loads never makes the epilogue wait on the
// prefetched chunks behind it).
template <int RT, typename XT>
struct StreamSlot {
  typename StreamVec<XT>::type a, b;
  float fj[4][RT];
  float sq[RT];
};

template <int PREC, typename XT, int RT, int EPI>
__device__ __forceinline__ void stream_finish_tile(const StreamArgs& args, int64_t tile, int c, int g,
                                                   const StreamSlot<RT, XT>& sl, const float wr[RT], floatx4 d[RT],
                                                   floatx4 acc[RT], const float* FjT, int LDJ) {
  const int R = args.R;
  const int64_t row0 = tile * 16;
  if constexpr (EPI == SE_REDQG) {
    // acc[nb][0]: this lane's part of sum_rows F_j[j, r] D[row, r] over the
    // q's tiles so far (rows 4 g .. 4 g + 3 of each tile)
    const int64_t tpq = args.J / 16;
    const int64_t q = tile / tpq;
    const int j0 = (int)(row0 - q * args.J);
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const floatx4 f = *(const floatx4*)(FjT + (nb * 16 + c) * LDJ + j0 + 4 * g);
      float t = acc[nb][0];
#pragma unroll
      for (int e = 0; e < 4; ++e) t = fmaf(f[e], d[nb][e], t);
      acc[nb][0] = t;
    }
    if ((tile + 1) % tpq == 0) {
#pragma unroll
      for (int nb = 0; nb < RT; ++nb) {
        const int r = nb * 16 + c;
        float t = acc[nb][0];
        t += __shfl_xor(t, 16, 64);
        t += __shfl_xor(t, 32, 64);
        if (g == 0 && r < R) args.out[q * R + r] = sl.sq[nb] * t;
        acc[nb][0] = 0.0f;
      }
    }
    return;
  }
  // d[nb][e] = D[row0 + 4 g + e][nb * 16 + c]
  if constexpr (EPI == SE_ROWS) {
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const int r = nb * 16 + c;
      if (r < R)
#pragma unroll
        for (int e = 0; e < 4; ++e) args.out[(row0 + 4 * g + e) * R + r] = wr[nb] * d[nb][e];
    }
  } else if constexpr (EPI == SE_REDQ) {
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const int r = nb * 16 + c;
      float t = 0.0f;
#pragma unroll
      for (int e = 0; e < 4; ++e) t = fmaf(sl.fj[e][nb], d[nb][e], t);
      t += __shfl_xor(t, 16, 64);
      t += __shfl_xor(t, 32, 64);
      if (g == 0 && r < R) args.out[tile * R + r] = sl.sq[nb] * t;
    }
  } else {
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) acc[nb] += sl.sq[nb] * d[nb];
  }
}

// Chunks in flight per wave (measured: 5 gains nothing over 3 at K = 128
// and K = 512 for float tensors).  A bf16 tensor's chunk is half the
// bytes, so it keeps more chunks in flight for the same bytes per wave.
constexpr int kStreamDepth = 3;
constexpr int kStreamDepthBf16 = 5;

template <int PREC, typename XT, int RT, int EPI>
__global__ __launch_bounds__(64 * kStreamWaves) void mttkrp_stream_kernel(StreamArgs args) {
  constexpr int DEPTH = sizeof(XT) == 2 ? kStreamDepthBf16 : kStreamDepth;
  constexpr int RP = RT * 16;
  constexpr int NT = 64 * kStreamWaves;
  extern __shared__ __align__(16) unsigned char smem[];
  if (take_inv_job<NT>(args.inv, (char*)smem)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int R = args.R;
  const int K = (int)args.K;  // K % 64 == 0, K * RP fits LDS
  const int LDF = K + 4, LDB = K + 8;
  float* Ct = (float*)smem;
  __bf16* Ch = (__bf16*)smem;
  __bf16* Cl = Ch + RP * LDB;
  for (int idx = tid; idx < RP * K; idx += NT) {
    const int n = idx % RP, kk = idx / RP;
    const float v = (n < R) ? args.Fk[(int64_t)kk * R + n] : 0.0f;
    if constexpr (PREC == TLA_PREC_FP32) {
      Ct[n * LDF + kk] = v;
    } else {
      const __bf16 h = (__bf16)v;
      Ch[n * LDB + kk] = h;
      Cl[n * LDB + kk] = (__bf16)(v - (float)h);
    }
  }
  // SE_REDQG: F_j transposed after the F_k panel, [RP][J + 4] (16-byte rows,
  // zero past R)
  const int LDJ = (int)args.J + 4;
  float* FjT = (float*)(smem + (PREC == TLA_PREC_FP32 ? size_t(RP) * LDF * 4 : size_t(2) * RP * LDB * 2));
  if constexpr (EPI == SE_REDQG) {
    const int J = (int)args.J;
    for (int idx = tid; idx < RP * J; idx += NT) {
      const int n = idx % RP, j = idx / RP;
      FjT[n * LDJ + j] = (n < R) ? args.Fj[(int64_t)j * R + n] : 0.0f;
    }
  }
  float wr[RT];
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) {
    const int r = nb * 16 + c;
    wr[nb] = (r < R && args.weights) ? args.weights[r] : 1.0f;
  }
  __syncthreads();

  const int64_t W = (int64_t)(gridDim.x - (args.inv.R ? 1 : 0)) * kStreamWaves;
  const int64_t w = (int64_t)blockIdx.x * kStreamWaves + wave;
  const XT* X = (const XT*)args.X;
  const int64_t tpq = args.J / 16;
  floatx4 acc[RT];
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  // tiles this wave visits: w, w + W, ...; SE_REDQG: all tpq tiles of
  // q = w, w + W, ... in turn
  const int64_t nq = args.ntiles / tpq;
  const int64_t my_tiles = (EPI == SE_REDQG) ? ((w < nq) ? (nq - w + W - 1) / W * tpq : 0)
                                             : ((w < args.ntiles) ? (args.ntiles - w + W - 1) / W : 0);
  const int cpt = K / 64;  // 64-column chunks per tile
  const int64_t nch = my_tiles * cpt;
  auto next_tile = [&](int64_t t) -> int64_t {
    if constexpr (EPI == SE_REDQG) return ((t + 1) % tpq == 0) ? t + 1 + (W - 1) * tpq : t + 1;
    else return t + W;
  };
  const int64_t first_tile = (EPI == SE_REDQG) ? w * tpq : w;
  // chunk i covers tile w + (i / cpt) * W, columns (i % cpt) * 64 .. +63.
  // DEPTH-deep ring: chunk i + DEPTH - 1 is requested while chunk i is
  // multiplied.
  int64_t ld_tile = first_tile;
  int ld_kk = 0;
  auto issue = [&](StreamSlot<RT, XT>& sl) {
    load_pair<XT, EPI != SE_ROWS>(X + (ld_tile * 16 + c) * args.K + ld_kk + 8 * g, sl.a, sl.b);
    if (ld_kk + 64 == K) {
      if constexpr (EPI == SE_REDQG) {
        if ((ld_tile + 1) % tpq == 0) {  // the q's last tile stores its partial
          const int64_t q = ld_tile / tpq;
#pragma unroll
          for (int nb = 0; nb < RT; ++nb) {
            const int r = nb * 16 + c;
            sl.sq[nb] = (r < R) ? args.Sq[q * R + r] : 0.0f;
          }
        }
      } else if constexpr (EPI != SE_ROWS) {
        const int64_t q = ld_tile / tpq;
#pragma unroll
        for (int nb = 0; nb < RT; ++nb) {
          const int r = nb * 16 + c;
          sl.sq[nb] = (r < R) ? args.Sq[q * R + r] : 0.0f;
        }
        if constexpr (EPI == SE_REDQ) {
          const int64_t j0 = ld_tile * 16 - q * args.J;
#pragma unroll
          for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int nb = 0; nb < RT; ++nb) {
              const int r = nb * 16 + c;
              sl.fj[e][nb] = (r < R) ? args.Fj[(j0 + 4 * g + e) * R + r] : 0.0f;
            }
        }
      }
    }
    ld_kk += 64;
    if (ld_kk == K) {
      ld_kk = 0;
      ld_tile = next_tile(ld_tile);
    }
  };
  StreamSlot<RT, XT> ring[DEPTH];
#pragma unroll
  for (int s = 0; s < DEPTH - 1; ++s)
    if (s < nch) issue(ring[s]);
  int64_t cur_tile = first_tile;
  int cur_kk = 0;
  floatx4 d[RT];
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) d[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto consume = [&](const StreamSlot<RT, XT>& sl) {
    stream_step<PREC, RT>(sl.a, Ct, Ch, Cl, LDF, LDB, cur_kk, c, g, d);
    stream_step<PREC, RT>(sl.b, Ct, Ch, Cl, LDF, LDB, cur_kk + 32, c, g, d);
    cur_kk += 64;
    if (cur_kk == K) {
      stream_finish_tile<PREC, XT, RT, EPI>(args, cur_tile, c, g, sl, wr, d, acc, FjT, LDJ);
#pragma unroll
      for (int nb = 0; nb < RT; ++nb) d[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
      cur_kk = 0;
      cur_tile = next_tile(cur_tile);
    }
  };
  for (int64_t i = 0; i < nch; i += DEPTH) {
#pragma unroll
    for (int s = 0; s < DEPTH; ++s) {
      if (i + s >= nch) break;
      if (i + s + DEPTH - 1 < nch) issue(ring[(s + DEPTH - 1) % DEPTH]);
      consume(ring[s]);
    }
  }
  if constexpr (EPI == SE_ACCJ) {
    // this wave's output rows: 16 * (w % tpq) + 4 g + e
#pragma unroll
    for (int nb = 0; nb < RT; ++nb) {
      const int r = nb * 16 + c;
      if (r < R)
#pragma unroll
        for (int e = 0; e < 4; ++e) args.out[(w * 16 + 4 * g + e) * R + r] = acc[nb][e];
    }
  }
}

// out[j, r] = sum over waves w with w % tpq == j / 16 of partial[w, j % 16, r]
__global__ void accj_reduce_kernel(const float* __restrict__ partial, int64_t W, int64_t J, int R,
                                   float* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= J * R) return;
  const int r = (int)(e % R);
  const int64_t j = e / R;
  const int64_t tpq = J / 16;
  float acc = 0.0f;
  for (int64_t w = j / 16; w < W; w += tpq) acc += partial[(w * 16 + (j % 16)) * R + r];
  out[e] = acc;
}

// M[i, r] = sum_{a < Pa} sum_{u < U} partial[((a In + i) U + u) R + r]
// (u = (b, s): the tiles of one output row i for one a): first over a into
// tmp (In U R floats) when Pa > 1, then over u.  Both passes are wide
// strided reductions.
static size_t tiles_reduce_tmp_elems(int64_t Pa, int64_t In, int64_t U, int R) {
  return (Pa > 1 ? size_t(In * U * R) : 0) + reduce_tmp_elems(In, R, U, U * R, R);
}

static void tiles_reduce(const float* partial, int64_t Pa, int64_t In, int64_t U, int R, float* tmp, float* out,
                         hipStream_t st) {
  const float* src = partial;
  float* tmp2 = tmp;
  if (Pa > 1) {
    strided_reduce(partial, In, U * R, Pa, U * R, In * U * R, tmp, st);
    src = tmp;
    tmp2 = tmp + In * U * R;
  }
  strided_reduce_wide(src, In, R, U, U * R, R, out, tmp2, st);
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
// the same with non-temporal (streaming) loads
template <typename XT>
__device__ __forceinline__ floatx4 load_x4_nt(const XT* p) {
  if constexpr (sizeof(XT) == 2) return __builtin_convertvector(__builtin_nontemporal_load((const bf16x4*)p), floatx4);
  else return __builtin_nontemporal_load((const floatx4*)p);
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
        split_bf16(w8, wh[mb], wl[mb]);
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
          acc[t][mb] = mfma_bf16_16x16x32(wl[mb], xh, acc[t][mb]);
          if constexpr (PREC == TLA_PREC_BF16X3) acc[t][mb] = mfma_bf16_16x16x32(wh[mb], xl, acc[t][mb]);
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
// Persistent interleaved last-mode kernel.  A block of 8 waves covers up to
// 8 column tiles of 64 output columns (one column group) and
// rsplit = 8 / (waves per row) row sub-ranges of 32 rows; blocks of a column
// group take row visits of 32 * rsplit rows interleaved across the grid, so
// the grid sweeps the tensor front to back.  Each wave writes its own
// partial [slot][Kc][R], slot = block-in-group * rsplit + sub-range.
// ---------------------------------------------------------------------------
struct LastStreamArgs {
  const void* X;
  int64_t K, P, J;
  int R;
  const float* Fj;
  const float* Gm;
  float* partial;
  int wpr, rsplit;
  int64_t G;        // blocks per column group
  int64_t nvisits;  // ceil(P / (32 * rsplit))
  int ldT;          // > 0: F_j staged in LDS transposed, [RT * 16][ldT]
  InvJob inv;       // optional side job for the grid's last workgroup
};

// One visit's operands: 8 rows of this lane's column quad plus the Khatri-Rao
// row factors of those rows (requested together, see StreamSlot).
template <typename XT> struct LastVec { using type = floatx4; };
template <> struct LastVec<__bf16> { using type = bf16x4; };
template <int RT, typename XT>
struct LastSlot {
  typename LastVec<XT>::type x[8];  // as loaded (see StreamVec)
  float gm[RT];
  float fj[8][RT];
};

// 4 consecutive tensor values at p, as loaded, non-temporal
template <typename XT>
__device__ __forceinline__ typename LastVec<XT>::type load_raw4_nt(const XT* p) {
  return __builtin_nontemporal_load((const typename LastVec<XT>::type*)p);
}

template <typename XT, int RT>
__device__ __forceinline__ void last_load(const LastStreamArgs& args, const XT* X, int64_t col, bool colok,
                                          int64_t rowb, int64_t pe, int c, LastSlot<RT, XT>& sl) {
  const int R = args.R;
  const int64_t K = args.K, J = args.J;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int64_t row = rowb + u;
    sl.x[u] = typename LastVec<XT>::type{};
    if (row < pe && colok) sl.x[u] = load_raw4_nt<XT>(X + row * K + col);
  }
  // rows rowb + u share aa because J % 8 == 0 and rowb % 8 == 0
  const bool okrow = rowb < pe;
  const int64_t aa = okrow ? rowb / J : 0;
  const int64_t jj = rowb - aa * J;
#pragma unroll
  for (int mb = 0; mb < RT; ++mb) {
    const int r = mb * 16 + c;
    sl.gm[mb] = (okrow && r < R) ? args.Gm[aa * R + r] : 0.0f;
#pragma unroll
    for (int u = 0; u < 8; ++u) sl.fj[u][mb] = ((rowb + u) < pe && r < R) ? args.Fj[(jj + u) * R + r] : 0.0f;
  }
}

template <int PREC, int RT, typename XT>
__device__ __forceinline__ void last_visit(const LastSlot<RT, XT>& sl, floatx4 acc[4][RT]) {
  floatx8 w8[RT];
#pragma unroll
  for (int mb = 0; mb < RT; ++mb)
#pragma unroll
    for (int u = 0; u < 8; ++u) w8[mb][u] = sl.gm[mb] * sl.fj[u][mb];
  if constexpr (PREC == TLA_PREC_FP32) {
    // MFMA u contracts rows rowb(g) + u across the four lane groups
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mb = 0; mb < RT; ++mb) acc[t][mb] = mfma_f32_16x16x4(w8[mb][u], sl.x[u][t], acc[t][mb]);
  } else {
    bf16x8 wh[RT], wl[RT];
#pragma unroll
    for (int mb = 0; mb < RT; ++mb) {
      split_bf16(w8[mb], wh[mb], wl[mb]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      bf16x8 xh, xl;
      if constexpr (sizeof(XT) == 2) {
        // a bf16 copy: column t of the 8 rows, as stored
#pragma unroll
        for (int u = 0; u < 8; ++u) xh[u] = sl.x[u][t];
      } else {
        floatx8 x8;
#pragma unroll
        for (int u = 0; u < 8; ++u) x8[u] = sl.x[u][t];
        if constexpr (PREC == TLA_PREC_BF16X3) split_bf16(x8, xh, xl);
        else xh = __builtin_convertvector(x8, bf16x8);
      }
#pragma unroll
      for (int mb = 0; mb < RT; ++mb) {
        acc[t][mb] = mfma_bf16_16x16x32(wh[mb], xh, acc[t][mb]);
        acc[t][mb] = mfma_bf16_16x16x32(wl[mb], xh, acc[t][mb]);
        if constexpr (PREC == TLA_PREC_BF16X3) acc[t][mb] = mfma_bf16_16x16x32(wh[mb], xl, acc[t][mb]);
      }
    }
  }
}

// LDS variant: the slot holds only the tensor rows and the Gm row; the
// F_j rows are read from the transposed LDS copy at visit time, which keeps
// two slots and 4 rank tiles of accumulators within the register file.
template <int RT, typename XT>
struct LastSlotL {
  typename LastVec<XT>::type x[8];
  float gm[RT];
  int jj;
};

template <typename XT, int RT>
__device__ __forceinline__ void last_load(const LastStreamArgs& args, const XT* X, int64_t col, bool colok,
                                          int64_t rowb, int64_t pe, int c, LastSlotL<RT, XT>& sl) {
  const int R = args.R;
  const int64_t K = args.K, J = args.J;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int64_t row = rowb + u;
    sl.x[u] = typename LastVec<XT>::type{};
    if (row < pe && colok) sl.x[u] = load_raw4_nt<XT>(X + row * K + col);
  }
  // rows rowb + u share aa because J % 8 == 0 and rowb % 8 == 0; a row
  // block past the end reads F_j row 0 (its x and gm are zero)
  const bool okrow = rowb < pe;
  const int64_t aa = okrow ? rowb / J : 0;
  sl.jj = okrow ? (int)(rowb - aa * J) : 0;
#pragma unroll
  for (int mb = 0; mb < RT; ++mb) {
    const int r = mb * 16 + c;
    sl.gm[mb] = (okrow && r < R) ? args.Gm[aa * R + r] : 0.0f;
  }
}

template <int PREC, int RT, typename XT>
__device__ __forceinline__ void last_visit(const LastSlotL<RT, XT>& sl, const float* fjs, int ldT, int c,
                                           floatx4 acc[4][RT]) {
  LastSlot<RT, XT> t;
#pragma unroll
  for (int u = 0; u < 8; ++u) t.x[u] = sl.x[u];
#pragma unroll
  for (int mb = 0; mb < RT; ++mb) {
    t.gm[mb] = sl.gm[mb];
    const float* f = fjs + (mb * 16 + c) * ldT + sl.jj;
    const floatx4 f0 = *(const floatx4*)f;
    const floatx4 f1 = *(const floatx4*)(f + 4);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      t.fj[u][mb] = f0[u];
      t.fj[u + 4][mb] = f1[u];
    }
  }
  last_visit<PREC, RT, XT>(t, acc);
}

template <int PREC, typename XT, int RT, bool FJL>
__global__ __launch_bounds__(512) void mttkrp_last_stream_kernel(LastStreamArgs args) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int R = args.R;
  const int64_t K = args.K;
  const int64_t cgroup = blockIdx.x / args.G;
  const int64_t b = blockIdx.x - cgroup * args.G;
  extern __shared__ __align__(16) float fjs[];  // FJL: F_j transposed, [RT * 16][ldT]
  if (take_inv_job<512>(args.inv, (char*)fjs)) return;
  if constexpr (FJL) {
    const int64_t J = args.J;
    const int ldT = args.ldT;
    for (int64_t e = tid; e < J * RT * 16; e += 512) {
      const int64_t j = e / (RT * 16);
      const int r = (int)(e - j * (RT * 16));
      fjs[r * ldT + j] = (r < R) ? args.Fj[j * R + r] : 0.0f;
    }
    __syncthreads();
  }
  const int wc = wave % args.wpr, rs = wave / args.wpr;
  if (rs >= args.rsplit) return;  // only when wpr * rsplit < 8; no block-wide sync follows
  const int64_t ctile = cgroup * 8 + wc;
  const int64_t col = ctile * 64 + 4 * c;
  const bool colok = col < K;
  const XT* X = (const XT*)args.X;
  floatx4 acc[4][RT];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int mb = 0; mb < RT; ++mb) acc[t][mb] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int64_t V = 32 * (int64_t)args.rsplit;
  // visit v covers rows v * V + 32 * rs .. +31; lane rows p0 + 8 g + u.
  auto rowb_of = [&](int64_t v) { return v * V + 32 * rs + 8 * g; };
  auto pe_of = [&](int64_t v) { return min(args.P, v * V + 32 * rs + 32); };
  using Slot = typename std::conditional<FJL, LastSlotL<RT, XT>, LastSlot<RT, XT>>::type;
  const int ldT = args.ldT;
  Slot sa, sb;
  int64_t va = b;
  if (va < args.nvisits) last_load<XT, RT>(args, X, col, colok, rowb_of(va), pe_of(va), c, sa);
  for (; va < args.nvisits; va += 2 * args.G) {
    const int64_t vb = va + args.G;
    if (vb < args.nvisits) last_load<XT, RT>(args, X, col, colok, rowb_of(vb), pe_of(vb), c, sb);
    if constexpr (FJL) last_visit<PREC, RT, XT>(sa, fjs, ldT, c, acc);
    else last_visit<PREC, RT, XT>(sa, acc);
    if (vb >= args.nvisits) break;
    const int64_t vc = vb + args.G;
    if (vc < args.nvisits) last_load<XT, RT>(args, X, col, colok, rowb_of(vc), pe_of(vc), c, sa);
    if constexpr (FJL) last_visit<PREC, RT, XT>(sb, fjs, ldT, c, acc);
    else last_visit<PREC, RT, XT>(sb, acc);
  }
  // acc[t][mb][e] = out[col + t][mb * 16 + 4 g + e]
  const int64_t slot = b * args.rsplit + rs;
  float* dst = args.partial + slot * K * R;
  if (colok) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t k = col + t;
#pragma unroll
      for (int mb = 0; mb < RT; ++mb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = mb * 16 + 4 * g + e;
          if (r < R) dst[k * R + r] = acc[t][mb][e];
        }
    }
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
      if (args.fac[m]) v *= args.fac[m][idx * args.R + r];
    }
  }
  gm[e] = v;
}

// ---------------------------------------------------------------------------
// Host planning
// ---------------------------------------------------------------------------
enum class Path { Generic, Inner, Tiles, Last, Stream, LastStream };

struct Plan {
  Path path;
  int RT;
  // stream
  StreamArgs st;
  int sepi;
  int64_t s_W;
  LastStreamArgs ls;
  // tiles
  TilesArgs tl;
  int epi;
  int64_t t_Pa, t_In, t_Pb;
  int t_S;
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
  size_t red_off;  // workspace offset (bytes) of the reductions' temporary storage
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
// reserve_block: leave one workgroup slot of the 256 free for a side job
// (InvJob) in the persistent kernels.
static int make_plan(int x_dtype, int ndim, const int64_t* shape, int R, int kb, int ke,
                     int precision, Plan* p, bool reserve_block = false) {
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "tensor order %d not in [2, %d]", ndim, kMaxOrder);
  TLA_REQUIRE(R >= 1, "rank must be >= 1");
  TLA_REQUIRE(0 <= kb && kb < ke && ke <= ndim, "bad kept-mode range [%d, %d)", kb, ke);
  for (int m = 0; m < ndim; ++m) TLA_REQUIRE(shape[m] >= 1, "empty mode %d", m);
  *p = Plan{};
  const bool standard = (ke == kb + 1);
  // the MFMA paths size their grids in 32-bit counts of 16-row / 64-column
  // tiles; beyond that a standard MTTKRP takes the generic kernel and a
  // partial one is refused
  constexpr int64_t kMaxTiles = 0x7fffffff;
  auto too_big = [&](int64_t tiles) -> int {
    TLA_REQUIRE(standard, "partial contraction: %lld tiles exceed the grid limit", (long long)tiles);
    TLA_REQUIRE(x_dtype != TLA_BF16, "bf16 tensor copy: %lld tiles exceed the grid limit", (long long)tiles);
    p->path = Path::Generic;
    p->ws_bytes = 0;
    return TLA_OK;
  };
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
    // Row-tile kernel when mode N-2 is contracted (rows reduce into the
    // output) or when every non-contiguous mode is kept (rows are the output).
    const int64_t J = shape[ndim - 2];
    int64_t Prow = 1;
    for (int m = 0; m < ndim - 1; ++m) Prow *= shape[m];
    if ((Prow + 15) / 16 > kMaxTiles || (In + 15) / 16 > kMaxTiles) return too_big(std::max(Prow, In) / 16);
    {
      // Persistent interleaved kernel when F_k fits LDS whole.
      const size_t lds = (precision == TLA_PREC_FP32) ? size_t(RT * 16) * (K + 4) * 4
                                                       : size_t(2) * RT * 16 * (K + 8) * 2;
      const bool s_rows = (kb == 0 && ke == ndim - 1);
      const bool s_redq = (ke <= ndim - 2 && J % 16 == 0);
      const bool s_accj = (standard && kb == ndim - 2 && J % 16 == 0);
      const bool off = getenv("TLA_NO_STREAM") != nullptr;
      if (!off && K % 64 == 0 && Prow % 16 == 0 && lds <= 80 * 1024 && (s_rows || s_redq || s_accj)) {
        p->path = Path::Stream;
        p->sepi = s_rows ? SE_ROWS : (s_redq ? SE_REDQ : SE_ACCJ);
        StreamArgs& s = p->st;
        s.K = K; s.P = Prow; s.J = J; s.R = R;
        s.ntiles = Prow / 16;
        const int64_t tpq = J / 16;
        constexpr int NW = kStreamWaves;
        int64_t blocks = reserve_block ? 255 : 256;
        if (const char* e = getenv("TLA_STREAM_BLOCKS")) blocks = std::max(1, atoi(e));
        blocks = std::min<int64_t>(blocks, (s.ntiles + NW - 1) / NW);
        int64_t Wv = blocks * NW;
        if (p->sepi == SE_ACCJ) {
          int64_t L = tpq;
          while (L % NW) L += tpq;  // lcm(waves per block, tpq)
          Wv = std::max<int64_t>(L, (Wv + L / 2) / L * L);
          blocks = Wv / NW;
        }
        p->s_W = Wv;
        // one wave per q (all its tiles in a row) when every wave gets at
        // least two values of q and F_j fits LDS beside the F_k panel
        const size_t fjt = size_t(RT * 16) * (J + 4) * 4;
        if (p->sepi == SE_REDQ && s.ntiles / tpq >= 2 * Wv && lds + fjt <= 150 * 1024)
          p->sepi = SE_REDQG;
        if (p->sepi == SE_ACCJ) {
          s.na = ndim - 2; s.nb = 0; s.In = 1; s.Pb = 1;
          for (int m = 0; m < ndim - 2; ++m) s.asz[m] = shape[m];
        } else {
          int64_t Pbr = 1;
          for (int m = ke; m < ndim - 2; ++m) Pbr *= shape[m];
          s.na = kb;
          s.nb = s_rows ? 0 : (ndim - 2 - ke);
          for (int m = 0; m < kb; ++m) s.asz[m] = shape[m];
          for (int m = ke; m < ndim - 2; ++m) s.bsz[m - ke] = shape[m];
          s.In = In;
          s.Pb = s_rows ? 1 : Pbr;
          p->t_Pa = Pa; p->t_In = In; p->t_Pb = Pbr; p->t_S = (int)tpq;
        }
        p->grid = dim3((unsigned)blocks, 1, 1);
        p->smem = lds + (p->sepi == SE_REDQG ? fjt : 0);
        p->partial_elems = (p->sepi == SE_ROWS)    ? 0
                           : (p->sepi == SE_REDQ)  ? size_t(s.ntiles) * R
                           : (p->sepi == SE_REDQG) ? size_t(s.ntiles / tpq) * R
                                                   : size_t(Wv) * 16 * R;
        // block scales Sq[q, r] = w[r] * prod of the factor rows of q's digits
        // (the output digit excluded), one row per 16 * tpq-row block
        GmArgs& sq = p->gm;
        sq = GmArgs{};
        sq.R = R;
        if (p->sepi == SE_ACCJ) {
          sq.n = ndim - 2;
          for (int m = 0; m < ndim - 2; ++m) sq.sz[m] = shape[m];
        } else if (p->sepi == SE_REDQ || p->sepi == SE_REDQG) {
          int n = 0;
          for (int m = 0; m < kb; ++m) sq.sz[n++] = shape[m];
          sq.sz[n++] = In;
          for (int m = ke; m < ndim - 2; ++m) sq.sz[n++] = shape[m];
          sq.n = n;
        }
        p->gm_rows = (p->sepi == SE_ROWS) ? 0 : s.ntiles / tpq;
        sq.rows = p->gm_rows;
        p->ws_bytes = align256(size_t(p->gm_rows) * R * 4) + align256(p->partial_elems * 4);
        p->red_off = p->ws_bytes;
        if (p->sepi == SE_REDQ) p->ws_bytes += align256(tiles_reduce_tmp_elems(Pa, In, p->t_Pb * tpq, R) * 4);
        if (p->sepi == SE_REDQG) p->ws_bytes += align256(tiles_reduce_tmp_elems(Pa, In, p->t_Pb, R) * 4);
        return TLA_OK;
      }
    }
    const bool rows_epi = (kb == 0 && ke == ndim - 1 && Prow % 16 == 0);
    const bool reduce_epi = (ke <= ndim - 2 && J % 16 == 0);
    if (rows_epi || reduce_epi) {
      p->path = Path::Tiles;
      p->epi = rows_epi ? EPI_ROWS : EPI_REDUCE;
      TilesArgs& t = p->tl;
      t.K = K; t.P = Prow; t.R = R;
      t.J = rows_epi ? Prow : J;
      int tpu = 4;
      if (const char* e = getenv("TLA_TILES_TPU")) tpu = std::max(1, std::min(4, atoi(e)));
      while (tpu > 1 && (t.J % (16 * tpu)) != 0) tpu >>= 1;
      t.tpu = tpu;
      t.n_units = Prow / (16 * tpu);
      int panel = 128;
      if (const char* e = getenv("TLA_TILES_PANEL")) panel = std::max(32, atoi(e) / 32 * 32);
      t.panel = (int)std::min<int64_t>(K, panel);
      // q = (a, i, b) with a = modes < kb, i = kept block, b = modes ke..N-3
      int64_t Pbr = 1;
      for (int m = ke; m < ndim - 2; ++m) Pbr *= shape[m];
      t.na = kb;
      t.nb = rows_epi ? 0 : (ndim - 2 - ke);
      for (int m = 0; m < kb; ++m) t.asz[m] = shape[m];
      for (int m = ke; m < ndim - 2; ++m) t.bsz[m - ke] = shape[m];
      t.In = In;
      t.Pb = rows_epi ? 1 : Pbr;
      p->t_Pa = Pa; p->t_In = In; p->t_Pb = Pbr;
      p->t_S = rows_epi ? 1 : (int)(J / (16 * tpu));
      p->grid = dim3((unsigned)ceil_div(t.n_units, 4), 1, 1);
      p->smem = (precision == TLA_PREC_FP32) ? size_t(RT * 16) * inner_pitch_f32(t.panel) * 4
                                             : size_t(2) * RT * 16 * inner_pitch_bf16(t.panel) * 2;
      p->partial_elems = rows_epi ? 0 : size_t(t.n_units) * R;
      p->ws_bytes = align256(p->partial_elems * 4);
      p->red_off = p->ws_bytes;
      if (!rows_epi) p->ws_bytes += align256(tiles_reduce_tmp_elems(Pa, In, Pbr * p->t_S, R) * 4);
      return TLA_OK;
    }
    // more K panels than a grid dimension holds: the generic kernel
    if ((K + 127) / 128 > 65535) return too_big((K + 127) / 128);
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
    p->red_off = p->ws_bytes;
    p->ws_bytes += align256(reduce_tmp_elems(1, In * R, p->splits, 0, In * R) * 4);
    return TLA_OK;
  }
  // Output block is a suffix including the contiguous mode: last kernel.
  int64_t Kc = 1;
  for (int m = kb; m < ndim; ++m) Kc *= shape[m];
  int64_t P = 1;
  for (int m = 0; m < kb; ++m) P *= shape[m];
  p->out_rows = Kc;
  if ((Kc + 63) / 64 > kMaxTiles / 8 && mfma) return too_big((Kc + 63) / 64);
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
  GmArgs& gm = p->gm;
  gm.n = kb - 1;
  gm.R = R;
  int64_t gm_rows = 1;
  for (int m = 0; m < kb - 1; ++m) { gm.sz[m] = shape[m]; gm_rows *= shape[m]; }
  gm.rows = gm_rows;
  p->gm_rows = gm_rows;
  if (getenv("TLA_NO_STREAM") == nullptr && J % 8 == 0) {
    p->path = Path::LastStream;
    LastStreamArgs& s = p->ls;
    s.K = Kc; s.P = P; s.J = J; s.R = R;
    const int64_t ktiles = ceil_div(Kc, 64);
    int wpr = 1;
    while (wpr < 8 && wpr < ktiles) wpr <<= 1;
    s.wpr = wpr;
    s.rsplit = 8 / wpr;
    const int64_t cgroups = (wpr == 8) ? ceil_div(ktiles, 8) : 1;
    s.nvisits = (P + 32 * s.rsplit - 1) / (32 * s.rsplit);
    int64_t target = reserve_block ? 255 : 256;
    if (const char* e = getenv("TLA_LAST_STREAM_BLOCKS")) target = std::max(1, atoi(e));
    int64_t G = std::max<int64_t>(1, target / cgroups);
    G = std::min<int64_t>(G, s.nvisits);
    s.G = G;
    // F_j in LDS (transposed, row stride J + 4: 16-byte aligned rows that
    // start 4 banks apart) when it fits; otherwise F_j rows ride in the slot
    const size_t fj_lds = size_t(RT) * 16 * (J + 4) * 4;
    const int fj_mode = getenv("TLA_LAST_FJ") ? atoi(getenv("TLA_LAST_FJ")) : -1;
    const bool use_lds = fj_mode == 1 || (fj_mode != 0 && RT >= 3);
    s.ldT = (use_lds && fj_lds <= 96 * 1024) ? (int)(J + 4) : 0;
    p->smem = s.ldT ? fj_lds : 0;
    p->grid = dim3((unsigned)(G * cgroups), 1, 1);
    p->splits = (int)(G * s.rsplit);
    p->partial_elems = size_t(p->splits) * Kc * R;
    p->ws_bytes = align256(size_t(gm_rows) * R * 4) + align256(p->partial_elems * 4);
    p->red_off = p->ws_bytes;
    p->ws_bytes += align256(reduce_tmp_elems(1, Kc * R, p->splits, 0, Kc * R) * 4);
    return TLA_OK;
  }
  p->path = Path::Last;
  LastArgs& l = p->last;
  l.K = Kc; l.P = P; l.J = J; l.R = R;
  const int ktiles = ceil_div(Kc, 64);
  const int step = precision == TLA_PREC_FP32 ? 4 : 32;
  int64_t target_blocks = 512;
  if (const char* e = getenv("TLA_LAST_BLOCKS")) target_blocks = std::max(1, atoi(e));
  int64_t S = std::max<int64_t>(1, (target_blocks + ktiles - 1) / ktiles);
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
  p->red_off = p->ws_bytes;
  p->ws_bytes += align256(reduce_tmp_elems(1, Kc * R, S, 0, Kc * R) * 4);
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

template <int PREC, typename XT, int EPI>
static int launch_tiles_epi(const Plan& p, hipStream_t st) {
  TilesArgs a = p.tl;
  switch (p.RT) {
    case 1: mttkrp_tiles_kernel<PREC, XT, 1, EPI><<<p.grid, 256, p.smem, st>>>(a); break;
    case 2: mttkrp_tiles_kernel<PREC, XT, 2, EPI><<<p.grid, 256, p.smem, st>>>(a); break;
    case 3: mttkrp_tiles_kernel<PREC, XT, 3, EPI><<<p.grid, 256, p.smem, st>>>(a); break;
    case 4: mttkrp_tiles_kernel<PREC, XT, 4, EPI><<<p.grid, 256, p.smem, st>>>(a); break;
    default:
      set_error("rank tile %d unsupported", p.RT);
      return TLA_ERR_UNSUPPORTED;
  }
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

// Kernels staging more than 64 KB of dynamic LDS must opt in (gfx950 has
// 160 KB per workgroup).
template <typename F>
static int allow_lds(F* fn, size_t bytes) {
  if (bytes > 64 * 1024)
    TLA_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return TLA_OK;
}

template <int PREC, typename XT, int RT, int EPI>
static int launch_stream_rt(const Plan& p, hipStream_t st) {
  auto* fn = &mttkrp_stream_kernel<PREC, XT, RT, EPI>;
  const size_t lds = p.st.inv.R ? std::max(p.smem, inv_job_lds_bytes(p.st.inv.R)) : p.smem;
  int rc = allow_lds(fn, lds);
  if (rc) return rc;
  fn<<<p.grid, 64 * kStreamWaves, lds, st>>>(p.st);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

template <int PREC, typename XT, int EPI>
static int launch_stream_epi(const Plan& p, hipStream_t st) {
  switch (p.RT) {
    case 1: return launch_stream_rt<PREC, XT, 1, EPI>(p, st);
    case 2: return launch_stream_rt<PREC, XT, 2, EPI>(p, st);
    case 3: return launch_stream_rt<PREC, XT, 3, EPI>(p, st);
    case 4: return launch_stream_rt<PREC, XT, 4, EPI>(p, st);
    default:
      set_error("rank tile %d unsupported", p.RT);
      return TLA_ERR_UNSUPPORTED;
  }
}

template <int PREC, typename XT>
static int launch_stream(const Plan& p, hipStream_t st) {
  if (p.sepi == SE_ROWS) return launch_stream_epi<PREC, XT, SE_ROWS>(p, st);
  if (p.sepi == SE_REDQ) return launch_stream_epi<PREC, XT, SE_REDQ>(p, st);
  if (p.sepi == SE_REDQG) return launch_stream_epi<PREC, XT, SE_REDQG>(p, st);
  return launch_stream_epi<PREC, XT, SE_ACCJ>(p, st);
}

template <int PREC, typename XT, int RT>
static int launch_last_stream_rt(const Plan& p, hipStream_t st) {
  const LastStreamArgs a = p.ls;
  const size_t lds = a.inv.R ? std::max(p.smem, inv_job_lds_bytes(a.inv.R)) : p.smem;
  if (a.ldT > 0) {
    auto* fn = &mttkrp_last_stream_kernel<PREC, XT, RT, true>;
    const int rc = allow_lds(fn, lds);
    if (rc) return rc;
    fn<<<p.grid, 512, lds, st>>>(a);
  } else {
    auto* fn = &mttkrp_last_stream_kernel<PREC, XT, RT, false>;
    const int rc = allow_lds(fn, lds);
    if (rc) return rc;
    fn<<<p.grid, 512, lds, st>>>(a);
  }
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

template <int PREC, typename XT>
static int launch_last_stream(const Plan& p, hipStream_t st) {
  switch (p.RT) {
    case 1: return launch_last_stream_rt<PREC, XT, 1>(p, st);
    case 2: return launch_last_stream_rt<PREC, XT, 2>(p, st);
    case 3: return launch_last_stream_rt<PREC, XT, 3>(p, st);
    case 4: return launch_last_stream_rt<PREC, XT, 4>(p, st);
    default:
      set_error("rank tile %d unsupported", p.RT);
      return TLA_ERR_UNSUPPORTED;
  }
}

template <int PREC, typename XT>
static int launch_tiles(const Plan& p, hipStream_t st) {
  return p.epi == EPI_ROWS ? launch_tiles_epi<PREC, XT, EPI_ROWS>(p, st)
                           : launch_tiles_epi<PREC, XT, EPI_REDUCE>(p, st);
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
                    int precision, void* out, void* ws, size_t ws_bytes, hipStream_t st,
                    const InvJob* job = nullptr, bool* job_taken = nullptr) {
  Plan p = p0;
  if (job_taken) *job_taken = false;
  // A matrix's MTTKRP multiplies by the one other factor alone: khatri_rao
  // of a single matrix returns it without the weights
  // (tensorly/tenalg/core_tenalg/_khatri_rao.py:67-69).
  if (ndim == 2 && ke == kb + 1) weights = nullptr;
  // the side job rides along as one extra workgroup when the grid leaves a
  // slot of the 256 free (all workgroups resident at once)
  if (job && job->R > 0 && (p.path == Path::Stream || p.path == Path::LastStream) && p.grid.x + 1 <= 256) {
    if (p.path == Path::Stream) p.st.inv = *job;
    else p.ls.inv = *job;
    p.grid.x += 1;
    if (job_taken) *job_taken = true;
  }
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
  if (p.path == Path::Stream) {
    StreamArgs& s = p.st;
    s.X = X;
    s.weights = (const float*)weights;
    s.Fj = (const float*)factors[ndim - 2];
    s.Fk = (const float*)factors[ndim - 1];
    float* sqbuf = (float*)ws;
    float* part = (float*)((char*)ws + align256(size_t(p.gm_rows) * R * 4));
    const bool rows_out = p.sepi == SE_ROWS;
    if (!rows_out) {
      GmArgs g = p.gm;
      if (p.sepi == SE_ACCJ) {
        for (int m = 0; m < g.n; ++m) g.fac[m] = (const float*)factors[m];
      } else {
        int n = 0;
        for (int m = 0; m < kb; ++m) g.fac[n++] = (const float*)factors[m];
        g.fac[n++] = nullptr;  // the output digit
        for (int m = ke; m < ndim - 2; ++m) g.fac[n++] = (const float*)factors[m];
      }
      g.weights = (const float*)weights;
      const int64_t ne = p.gm_rows * R;
      build_gm_kernel<<<ceil_div(ne, 256), 256, 0, st>>>(g, sqbuf);
      TLA_LAUNCH_CHECK();
    }
    s.Sq = sqbuf;
    s.out = rows_out ? (float*)out : part;
    if (x_dtype == TLA_BF16) rc = launch_stream<TLA_PREC_BF16, __bf16>(p, st);
    else if (precision == TLA_PREC_FP32) rc = launch_stream<TLA_PREC_FP32, float>(p, st);
    else if (precision == TLA_PREC_BF16) rc = launch_stream<TLA_PREC_BF16, float>(p, st);
    else rc = launch_stream<TLA_PREC_BF16X3, float>(p, st);
    if (rc != TLA_OK || rows_out) return rc;
    if (p.sepi == SE_REDQ) {
      tiles_reduce(part, p.t_Pa, p.t_In, p.t_Pb * p.t_S, R, (float*)((char*)ws + p.red_off), (float*)out, st);
    } else if (p.sepi == SE_REDQG) {
      tiles_reduce(part, p.t_Pa, p.t_In, p.t_Pb, R, (float*)((char*)ws + p.red_off), (float*)out, st);
    } else {
      const int64_t n = s.J * R;
      (void)n;
      const int64_t tpq = s.J / 16;
      strided_reduce(part, tpq, 16 * R, p.s_W / tpq, 16 * R, tpq * 16 * R, (float*)out, st);
    }
    TLA_LAUNCH_CHECK();
    return TLA_OK;
  }
  if (p.path == Path::LastStream) {
    LastStreamArgs& s = p.ls;
    float* gm = (float*)ws;
    partial = (float*)((char*)ws + align256(size_t(p.gm_rows) * R * 4));
    GmArgs g = p.gm;
    for (int m = 0; m < g.n; ++m) g.fac[m] = (const float*)factors[m];
    g.weights = (const float*)weights;
    if (g.n == 1 && !weights) {
      gm = (float*)factors[0];  // the Khatri-Rao row factor is that factor itself
    } else {
      const int64_t ne = p.gm_rows * R;
      build_gm_kernel<<<ceil_div(ne, 256), 256, 0, st>>>(g, gm);
      TLA_LAUNCH_CHECK();
    }
    s.X = X;
    s.Fj = (const float*)factors[kb - 1];
    s.Gm = gm;
    s.partial = partial;
    if (x_dtype == TLA_BF16) rc = launch_last_stream<TLA_PREC_BF16, __bf16>(p, st);
    else if (precision == TLA_PREC_FP32) rc = launch_last_stream<TLA_PREC_FP32, float>(p, st);
    else if (precision == TLA_PREC_BF16) rc = launch_last_stream<TLA_PREC_BF16, float>(p, st);
    else rc = launch_last_stream<TLA_PREC_BF16X3, float>(p, st);
    if (rc != TLA_OK) return rc;
    const int64_t n = p.out_rows * R;
    strided_reduce_wide(partial, 1, n, p.splits, 0, n, (float*)out, (float*)((char*)ws + p.red_off), st);
    TLA_LAUNCH_CHECK();
    return TLA_OK;
  }
  if (p.path == Path::Tiles) {
    TilesArgs& t = p.tl;
    t.X = X;
    for (int m = 0; m < t.na; ++m) t.afac[m] = (const float*)factors[m];
    for (int m = 0; m < t.nb; ++m) t.bfac[m] = (const float*)factors[ke + m];
    t.weights = (const float*)weights;
    t.Fj = (const float*)factors[ndim - 2];
    t.Fk = (const float*)factors[ndim - 1];
    t.out = (p.epi == EPI_ROWS) ? (float*)out : (float*)ws;
    if (x_dtype == TLA_BF16) rc = launch_tiles<TLA_PREC_BF16, __bf16>(p, st);
    else if (precision == TLA_PREC_FP32) rc = launch_tiles<TLA_PREC_FP32, float>(p, st);
    else if (precision == TLA_PREC_BF16) rc = launch_tiles<TLA_PREC_BF16, float>(p, st);
    else rc = launch_tiles<TLA_PREC_BF16X3, float>(p, st);
    if (rc != TLA_OK || p.epi == EPI_ROWS) return rc;
    tiles_reduce((const float*)ws, p.t_Pa, p.t_In, p.t_Pb * p.t_S, R, (float*)((char*)ws + p.red_off), (float*)out,
                 st);
    TLA_LAUNCH_CHECK();
    return TLA_OK;
  }
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
  strided_reduce_wide(partial, 1, n, p.splits, 0, n, (float*)out, (float*)((char*)ws + p.red_off), st);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

}  // namespace tla

using namespace tla;

int tla::mttkrp_launch(int x_dtype, const void* X, int ndim, const int64_t* shape, const void* const* factors,
                       const void* weights, int rank, int kb, int ke, int precision, void* out, void* ws,
                       size_t ws_bytes, const InvJob* job, bool* job_taken, hipStream_t st) {
  if (job_taken) *job_taken = false;
  TLA_REQUIRE(X && out && factors, "NULL pointer argument");
  if (x_dtype == TLA_F64 && precision != TLA_PREC_VALU) precision = TLA_PREC_VALU;
  Plan p;
  int rc = make_plan(x_dtype, ndim, shape, rank, kb, ke, precision, &p, job != nullptr && job->R > 0);
  if (rc != TLA_OK) return rc;
  TLA_REQUIRE(ke == kb + 1 || p.path != Path::Generic, "partial MTTKRP needs an MFMA-eligible shape");
  if (job) TLA_REQUIRE(job->R >= 0 && job->R <= kMaxRank && job->grams && job->vi && job->vist, "bad side job");
  return run_plan(p, x_dtype, X, ndim, shape, factors, weights, rank, kb, ke, precision, out, ws, ws_bytes, st, job,
                  job_taken);
}

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

extern "C" int tla_mttkrp_partial_supported(int x_dtype, int ndim, const int64_t* shape, int rank, int keep_begin,
                                            int keep_end, int precision) {
  if (!(keep_begin == 0 || keep_end == ndim) || (keep_begin == 0 && keep_end == ndim)) return 0;
  Plan p;
  if (make_plan(x_dtype, ndim, shape, rank, keep_begin, keep_end, precision, &p) != TLA_OK) return 0;
  return p.path != Path::Generic ? 1 : 0;
}
