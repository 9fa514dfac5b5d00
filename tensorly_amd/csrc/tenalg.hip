// Khatri-Rao product, n-mode product (TTM / TTV) and the rank-diagonal
// contraction used by the dimension-tree MTTKRP.
//
// mode_dot has two MFMA kernels for the Tucker chain (FP32 MFMA, BF16 or
// split BF16x3, FP32 accumulation):
//  * ttm_rows   -- contracting the contiguous (last) mode:
//                  out[p, j] = sum_k X[p, k] U'[j, k], 16 rows per wave, the
//                  U' panel staged transposed in LDS;
//  * ttm_cols   -- contracting an inner mode: out[a, j, b] =
//                  sum_m U'[j, m] X[a, m, b] with X as the MFMA B operand
//                  (16-byte loads of 4 consecutive b per lane, 256 bytes per
//                  row per wave) and the m-range split over the 4 waves of a
//                  block, reduced in LDS in a fixed order.
// and a vector-ALU kernel for everything else (vectors, f64, J > 64, odd
// sizes).
#include <algorithm>

#include "common.h"

namespace tla {

// ---------------------------------------------------------------------------
// Khatri-Rao
// ---------------------------------------------------------------------------
struct KrArgs {
  int n, R;
  int64_t rows[kMaxOrder];
  const void* mats[kMaxOrder];
  const void* weights;
  const void* mask;
  int64_t total;
};

template <typename T>
__global__ void khatri_rao_kernel(KrArgs a, T* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.total * a.R) return;
  const int r = (int)(e % a.R);
  int64_t q = e / a.R;
  const int64_t q0 = q;
  const T w = a.weights ? ((const T*)a.weights)[r] : (T)1;
  int64_t idx[kMaxOrder];
#pragma unroll
  for (int m = kMaxOrder - 1; m >= 0; --m) {
    if (m < a.n) {
      idx[m] = q % a.rows[m];
      q /= a.rows[m];
    }
  }
  // (w * A_0) (x) A_1 (x) ... in that rounding order
  T res = ((const T*)a.mats[0])[idx[0] * a.R + r] * w;
#pragma unroll
  for (int m = 1; m < kMaxOrder; ++m)
    if (m < a.n) res = res * ((const T*)a.mats[m])[idx[m] * a.R + r];
  if (a.mask) res = res * ((const T*)a.mask)[q0];
  out[e] = res;
}

// ---------------------------------------------------------------------------
// mode_dot, vector-ALU: out[a, j, b] = sum_m Uj(j, m) X[a, m, b]
// Uj(j, m) = U[j * ldj + m * ldm]
// ---------------------------------------------------------------------------
template <typename T>
__global__ void ttm_generic_kernel(const T* __restrict__ X, int64_t Pa, int64_t Im, int64_t Pb,
                                   const T* __restrict__ U, int64_t J, int64_t ldj, int64_t ldm,
                                   T* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Pa * J * Pb) return;
  const int64_t b = e % Pb;
  const int64_t t = e / Pb;
  const int64_t j = t % J, a = t / J;
  const T* x = X + a * Im * Pb + b;
  const T* u = U + j * ldj;
  T acc = 0;
  for (int64_t m = 0; m < Im; ++m) acc += u[m * ldm] * x[m * Pb];
  out[e] = acc;
}

// ---------------------------------------------------------------------------
// ttm_rows: out[p, j] = sum_k X[p, k] Ut[k, j]   (Ut = K x J, row-major view
// given by ld strides), J <= 64.
// ---------------------------------------------------------------------------
struct TtmRowsArgs {
  const float* X;
  int64_t P, K;
  const float* U;
  int64_t ldj, ldk;  // U'(j, k) = U[j * ldj + k * ldk]
  int J;
  float* out;
  int panel;
};

template <int PREC, int RT>
__global__ __launch_bounds__(256) void ttm_rows_kernel(TtmRowsArgs a) {
  constexpr int RP = RT * 16;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int LDF = a.panel + 4, LDB = a.panel + 8;
  float* Ct = (float*)smem;
  __bf16* Ch = (__bf16*)smem;
  __bf16* Cl = Ch + RP * LDB;
  const int64_t p = ((int64_t)blockIdx.x * 4 + wave) * 16 + c;
  const int64_t rp = p < a.P ? p : a.P - 1;
  const float* row = a.X + rp * a.K + 8 * g;
  floatx4 d[RT];
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) d[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int64_t kb = 0; kb < a.K; kb += a.panel) {
    const int klen = (int)min((int64_t)a.panel, a.K - kb);
    __syncthreads();
    for (int idx = tid; idx < RP * klen; idx += 256) {
      const int n = idx % RP, kk = idx / RP;
      const float v = (n < a.J) ? a.U[n * a.ldj + (kb + kk) * a.ldk] : 0.0f;
      if constexpr (PREC == TLA_PREC_FP32) {
        Ct[n * LDF + kk] = v;
      } else {
        const __bf16 h = (__bf16)v;
        Ch[n * LDB + kk] = h;
        Cl[n * LDB + kk] = (__bf16)(v - (float)h);
      }
    }
    __syncthreads();
    for (int kk = 0; kk < klen; kk += 32) {
      const floatx4 x0 = *(const floatx4*)(row + kb + kk);
      const floatx4 x1 = *(const floatx4*)(row + kb + kk + 4);
      floatx8 av;
      av[0] = x0[0]; av[1] = x0[1]; av[2] = x0[2]; av[3] = x0[3];
      av[4] = x1[0]; av[5] = x1[1]; av[6] = x1[2]; av[7] = x1[3];
      if constexpr (PREC == TLA_PREC_FP32) {
#pragma unroll
        for (int nb = 0; nb < RT; ++nb) {
          const float* bp = Ct + (nb * 16 + c) * LDF + kk + 8 * g;
          const floatx4 b0 = *(const floatx4*)bp;
          const floatx4 b1 = *(const floatx4*)(bp + 4);
#pragma unroll
          for (int v = 0; v < 4; ++v) d[nb] = mfma_f32_16x16x4(av[v], b0[v], d[nb]);
#pragma unroll
          for (int v = 0; v < 4; ++v) d[nb] = mfma_f32_16x16x4(av[4 + v], b1[v], d[nb]);
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
  }
  // d[nb][e] = out[tile row 4g + e][nb*16 + c]
  const int64_t tile0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
#pragma unroll
  for (int nb = 0; nb < RT; ++nb) {
    const int j = nb * 16 + c;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t pr = tile0 + 4 * g + e;
      if (pr < a.P && j < a.J) a.out[pr * a.J + j] = d[nb][e];
    }
  }
}

// ---------------------------------------------------------------------------
// ttm_cols: out[a, j, b] = sum_m U'(j, m) X[a, m, b], J <= 64, Pb % 4 == 0.
// grid.x = 64-column tiles of b, grid.y = a; 4 waves split the m range.
// ---------------------------------------------------------------------------
struct TtmColsArgs {
  const float* X;
  int64_t Pa, Im, Pb;
  const float* U;
  int64_t ldj, ldm;
  int J;
  float* out;
};

template <int PREC, int RT>
__global__ __launch_bounds__(256) void ttm_cols_kernel(TtmColsArgs a) {
  constexpr int RP = RT * 16;
  constexpr int STEP = (PREC == TLA_PREC_FP32) ? 4 : 32;
  __shared__ float red[64 * RP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int64_t col = (int64_t)blockIdx.x * 64 + 4 * c;
  const bool colok = col < a.Pb;
  const int64_t ia = blockIdx.y;
  const float* Xa = a.X + ia * a.Im * a.Pb;
  // m-range of this wave (multiple of STEP except the end)
  const int64_t per = ((a.Im + 4 * STEP - 1) / (4 * STEP)) * STEP;
  const int64_t mb0 = wave * per;
  const int64_t me = min(a.Im, mb0 + per);
  floatx4 acc[4][RT];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int mb = 0; mb < RT; ++mb) acc[t][mb] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int64_t m0 = mb0; m0 < me; m0 += STEP) {
    if constexpr (PREC == TLA_PREC_FP32) {
      const int64_t m = m0 + g;
      const bool ok = m < me;
      floatx4 xv = floatx4{0.f, 0.f, 0.f, 0.f};
      if (ok && colok) xv = *(const floatx4*)(Xa + m * a.Pb + col);
      float uv[RT];
#pragma unroll
      for (int mb = 0; mb < RT; ++mb) {
        const int j = mb * 16 + c;
        uv[mb] = (ok && j < a.J) ? a.U[j * a.ldj + m * a.ldm] : 0.0f;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int mb = 0; mb < RT; ++mb) acc[t][mb] = mfma_f32_16x16x4(uv[mb], xv[t], acc[t][mb]);
    } else {
      floatx4 xv[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int64_t m = m0 + 8 * g + v;
        xv[v] = floatx4{0.f, 0.f, 0.f, 0.f};
        if (m < me && colok) xv[v] = *(const floatx4*)(Xa + m * a.Pb + col);
      }
      bf16x8 uh[RT], ul[RT];
#pragma unroll
      for (int mb = 0; mb < RT; ++mb) {
        const int j = mb * 16 + c;
        floatx8 u8;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          const int64_t m = m0 + 8 * g + v;
          u8[v] = (m < me && j < a.J) ? a.U[j * a.ldj + m * a.ldm] : 0.0f;
        }
        split_bf16(u8, uh[mb], ul[mb]);
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
          acc[t][mb] = mfma_bf16_16x16x32(uh[mb], xh, acc[t][mb]);
          acc[t][mb] = mfma_bf16_16x16x32(ul[mb], xh, acc[t][mb]);
          if constexpr (PREC == TLA_PREC_BF16X3) acc[t][mb] = mfma_bf16_16x16x32(uh[mb], xl, acc[t][mb]);
        }
      }
    }
  }
  for (int w = 0; w < 4; ++w) {
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
  // out[a, j, b]: b fastest -- write 64 consecutive b for each j
  for (int idx = tid; idx < 64 * RP; idx += 256) {
    const int j = idx / 64, bc = idx % 64;
    const int64_t b = (int64_t)blockIdx.x * 64 + bc;
    if (b < a.Pb && j < a.J) a.out[(ia * a.J + j) * a.Pb + b] = red[bc * RP + j];
  }
}

// ---------------------------------------------------------------------------
// Rank-diagonal contraction of Y (I x J x R)
// ---------------------------------------------------------------------------
// axis 1: out[i, r] = sum_j Y[i, j, r] F[j, r]  -- one block per i; lane
// group jl of L = 256 / R sums j = jl, jl + L, ... with four independent
// partial sums, then the L groups are combined in order.
__global__ __launch_bounds__(256) void rankdiag_axis1_kernel(const float* __restrict__ Y, int64_t J, int R,
                                                             const float* __restrict__ F, float* __restrict__ out) {
  const int64_t i = blockIdx.x;
  __shared__ float red[256];
  const int L = 256 / R;
  const int r = threadIdx.x % R, jl = threadIdx.x / R;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (jl < L) {
    const float* y = Y + i * J * R + r;
    const float* f = F + r;
    int64_t j = jl;
    for (; j + 3 * L < J; j += 4 * L) {
      s0 = fmaf(y[j * R], f[j * R], s0);
      s1 = fmaf(y[(j + L) * R], f[(j + L) * R], s1);
      s2 = fmaf(y[(j + 2 * L) * R], f[(j + 2 * L) * R], s2);
      s3 = fmaf(y[(j + 3 * L) * R], f[(j + 3 * L) * R], s3);
    }
    for (; j < J; j += L) s0 = fmaf(y[j * R], f[j * R], s0);
  }
  red[threadIdx.x] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (threadIdx.x < R) {
    float t = 0.0f;
    for (int k = 0; k < L; ++k) t += red[k * R + threadIdx.x];
    out[i * R + threadIdx.x] = t;
  }
}

// axis 0: out[j, r] = sum_i Y[i, j, r] F[i, r] -- block = 4 i-lanes x 64
// consecutive (j, r) elements, every lane over a quarter of i, combined in
// order (single launch, no workspace).
__global__ __launch_bounds__(256) void rankdiag_axis0_kernel(const float* __restrict__ Y, int64_t I, int64_t J, int R,
                                                             const float* __restrict__ F, float* __restrict__ out) {
  __shared__ float red[4][64];
  const int el = threadIdx.x & 63, il = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + el;
  const int64_t JR = J * R;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (e < JR) {
    const int r = (int)(e % R);
    int64_t i = il;
    for (; i + 12 < I; i += 16) {
      s0 = fmaf(Y[i * JR + e], F[i * R + r], s0);
      s1 = fmaf(Y[(i + 4) * JR + e], F[(i + 4) * R + r], s1);
      s2 = fmaf(Y[(i + 8) * JR + e], F[(i + 8) * R + r], s2);
      s3 = fmaf(Y[(i + 12) * JR + e], F[(i + 12) * R + r], s3);
    }
    for (; i < I; i += 4) s0 = fmaf(Y[i * JR + e], F[i * R + r], s0);
  }
  red[il][el] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (il == 0 && e < JR) out[e] = (red[0][el] + red[1][el]) + (red[2][el] + red[3][el]);
}

// Vectorised forms for rank % 4 == 0 and 16-byte aligned operands.
// axis 1: one block per i; lane (row lane jl, column quad c) accumulates
// Y[i, j, 4c..4c+3] * F[j, 4c..4c+3] over rows j = jl, jl + L, ... with two
// independent float4 sums (loads of four rows in flight), then the L row
// lanes are summed in order.
__global__ __launch_bounds__(256) void rankdiag_axis1_v4_kernel(const float* __restrict__ Y, int64_t J, int R,
                                                                const float* __restrict__ F,
                                                                float* __restrict__ out) {
  const int64_t i = blockIdx.x;
  __shared__ float4 red[256];
  const int C = R >> 2;
  const int L = 256 / C;
  const int c = threadIdx.x % C, jl = threadIdx.x / C;
  float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
  if (jl < L) {
    const float4* y = reinterpret_cast<const float4*>(Y + i * J * R) + c;
    const float4* f = reinterpret_cast<const float4*>(F) + c;
    int64_t j = jl;
    for (; j + 3 * L < J; j += 4 * L) {
      const float4 y0 = y[j * C], y1 = y[(j + L) * C], y2 = y[(j + 2 * L) * C], y3 = y[(j + 3 * L) * C];
      const float4 f0 = f[j * C], f1 = f[(j + L) * C], f2 = f[(j + 2 * L) * C], f3 = f[(j + 3 * L) * C];
      s0.x = fmaf(y0.x, f0.x, s0.x); s0.y = fmaf(y0.y, f0.y, s0.y); s0.z = fmaf(y0.z, f0.z, s0.z); s0.w = fmaf(y0.w, f0.w, s0.w);
      s1.x = fmaf(y1.x, f1.x, s1.x); s1.y = fmaf(y1.y, f1.y, s1.y); s1.z = fmaf(y1.z, f1.z, s1.z); s1.w = fmaf(y1.w, f1.w, s1.w);
      s0.x = fmaf(y2.x, f2.x, s0.x); s0.y = fmaf(y2.y, f2.y, s0.y); s0.z = fmaf(y2.z, f2.z, s0.z); s0.w = fmaf(y2.w, f2.w, s0.w);
      s1.x = fmaf(y3.x, f3.x, s1.x); s1.y = fmaf(y3.y, f3.y, s1.y); s1.z = fmaf(y3.z, f3.z, s1.z); s1.w = fmaf(y3.w, f3.w, s1.w);
    }
    for (; j < J; j += L) {
      const float4 y0 = y[j * C], f0 = f[j * C];
      s0.x = fmaf(y0.x, f0.x, s0.x); s0.y = fmaf(y0.y, f0.y, s0.y); s0.z = fmaf(y0.z, f0.z, s0.z); s0.w = fmaf(y0.w, f0.w, s0.w);
    }
  }
  red[threadIdx.x] = make_float4(s0.x + s1.x, s0.y + s1.y, s0.z + s1.z, s0.w + s1.w);
  __syncthreads();
  if (threadIdx.x < R) {
    const int cc = threadIdx.x >> 2, comp = threadIdx.x & 3;
    float t = 0.0f;
    for (int k = 0; k < L; ++k) {
      const float4 v = red[k * C + cc];
      t += comp == 0 ? v.x : (comp == 1 ? v.y : (comp == 2 ? v.z : v.w));
    }
    out[i * R + threadIdx.x] = t;
  }
}

// axis 0, first stage: blockIdx.y = split s of the i range; block = 4 i-lanes
// x 64 float4 lanes over (j, r); partial[s][j * R + r] (second stage: a
// fixed-order strided reduction over the splits).
__global__ __launch_bounds__(256) void rankdiag_axis0_v4_kernel(const float* __restrict__ Y, int64_t I, int64_t J,
                                                                int R, int64_t i_per_split,
                                                                const float* __restrict__ F,
                                                                float* __restrict__ partial) {
  __shared__ float4 red[4][64];
  const int el = threadIdx.x & 63, il = threadIdx.x >> 6;
  const int64_t JR4 = J * R / 4;
  const int64_t e4 = (int64_t)blockIdx.x * 64 + el;
  const int64_t i0 = (int64_t)blockIdx.y * i_per_split;
  const int64_t i1 = min(I, i0 + i_per_split);
  float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
  if (e4 < JR4) {
    const int c = (int)(e4 % (R / 4));
    const float4* y = reinterpret_cast<const float4*>(Y) + e4;
    const float4* f = reinterpret_cast<const float4*>(F) + c;
    int64_t i = i0 + il;
    for (; i + 28 < i1; i += 32) {  // eight rows in flight per lane
      float4 yv[8], fv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        yv[u] = y[(i + 4 * u) * JR4];
        fv[u] = f[(i + 4 * u) * (R / 4)];
      }
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        s0.x = fmaf(yv[u].x, fv[u].x, s0.x); s0.y = fmaf(yv[u].y, fv[u].y, s0.y);
        s0.z = fmaf(yv[u].z, fv[u].z, s0.z); s0.w = fmaf(yv[u].w, fv[u].w, s0.w);
        s1.x = fmaf(yv[u + 1].x, fv[u + 1].x, s1.x); s1.y = fmaf(yv[u + 1].y, fv[u + 1].y, s1.y);
        s1.z = fmaf(yv[u + 1].z, fv[u + 1].z, s1.z); s1.w = fmaf(yv[u + 1].w, fv[u + 1].w, s1.w);
      }
    }
    for (; i < i1; i += 4) {
      const float4 y0 = y[i * JR4], f0 = f[i * (R / 4)];
      s0.x = fmaf(y0.x, f0.x, s0.x); s0.y = fmaf(y0.y, f0.y, s0.y); s0.z = fmaf(y0.z, f0.z, s0.z); s0.w = fmaf(y0.w, f0.w, s0.w);
    }
  }
  red[il][el] = make_float4(s0.x + s1.x, s0.y + s1.y, s0.z + s1.z, s0.w + s1.w);
  __syncthreads();
  if (il == 0 && e4 < JR4) {
    const float4 a = red[0][el], b = red[1][el], c2 = red[2][el], d = red[3][el];
    reinterpret_cast<float4*>(partial + (int64_t)blockIdx.y * J * R)[e4] =
        make_float4((a.x + b.x) + (c2.x + d.x), (a.y + b.y) + (c2.y + d.y), (a.z + b.z) + (c2.z + d.z),
                    (a.w + b.w) + (c2.w + d.w));
  }
}

static int64_t axis0_splits(int64_t I, int64_t J, int R) {
  const int64_t tiles = (J * R / 4 + 63) / 64;
  int64_t S = (1024 + tiles - 1) / tiles;
  S = std::max<int64_t>(1, std::min<int64_t>(S, I / 32));
  return std::min<int64_t>(S, 64);
}

__global__ void sum_splits_kernel(const float* __restrict__ partial, int64_t n, int splits,
                                  float* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += partial[(int64_t)k * n + e];
  out[e] = s;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <int PREC>
static int launch_ttm_rows(const TtmRowsArgs& a, hipStream_t st) {
  const int RT = (a.J + 15) / 16;
  const size_t smem = (PREC == TLA_PREC_FP32) ? size_t(RT * 16) * (a.panel + 4) * 4
                                              : size_t(2) * RT * 16 * (a.panel + 8) * 2;
  const dim3 grid((unsigned)ceil_div(a.P, 64));
  switch (RT) {
    case 1: ttm_rows_kernel<PREC, 1><<<grid, 256, smem, st>>>(a); break;
    case 2: ttm_rows_kernel<PREC, 2><<<grid, 256, smem, st>>>(a); break;
    case 3: ttm_rows_kernel<PREC, 3><<<grid, 256, smem, st>>>(a); break;
    default: ttm_rows_kernel<PREC, 4><<<grid, 256, smem, st>>>(a); break;
  }
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

template <int PREC>
static int launch_ttm_cols(const TtmColsArgs& a, hipStream_t st) {
  const int RT = (a.J + 15) / 16;
  const dim3 grid((unsigned)ceil_div(a.Pb, 64), (unsigned)a.Pa);
  switch (RT) {
    case 1: ttm_cols_kernel<PREC, 1><<<grid, 256, 0, st>>>(a); break;
    case 2: ttm_cols_kernel<PREC, 2><<<grid, 256, 0, st>>>(a); break;
    case 3: ttm_cols_kernel<PREC, 3><<<grid, 256, 0, st>>>(a); break;
    default: ttm_cols_kernel<PREC, 4><<<grid, 256, 0, st>>>(a); break;
  }
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}


// ---------------------------------------------------------------------------
// MTTKRP of a sparse tensor in coordinate (COO) form.  The nonzeros come
// ordered by their mode-n index (perm; row_ptr delimits each output row) and
// cut into segments of at most seg_len nonzeros within a row, so that a mode
// with few rows still spreads over many waves:
//   pass 1, one wave per segment: partial[j][r] = sum over the segment, in
//           order, of val * prod_{m != n} A_m[i_m, r];
//   pass 2, one wave per row: out[i][r] = w[r] * sum of the row's segment
//           partials in segment order.
// No atomics: repeatable bit for bit.
// ---------------------------------------------------------------------------
struct CooArgs {
  const int64_t* coords;  // [ndim][nnz]
  int64_t nnz;
  const void* vals;
  const int64_t* perm;
  const int64_t* row_ptr;      // rows + 1
  const int64_t* seg_row;      // n_seg
  const int64_t* seg_start;    // n_seg
  const int64_t* seg_row_ptr;  // rows + 1
  int64_t n_seg, seg_len;
  const void* factors[kMaxOrder];
  const void* weights;
  int ndim, mode, R;
  int64_t rows;
  void* partial;  // n_seg x R
  void* out;
};

template <typename T>
__global__ __launch_bounds__(256) void mttkrp_coo_segments_kernel(CooArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const T* vals = (const T*)a.vals;
  const int R = a.R;
  for (int64_t j = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); j < a.n_seg; j += nw) {
    const int64_t row = a.seg_row[j];
    const int64_t p0 = a.seg_start[j];
    const int64_t p1 = min(p0 + a.seg_len, a.row_ptr[row + 1]);
    if (lane < R) {
      T acc = (T)0;
      for (int64_t p = p0; p < p1; ++p) {
        const int64_t q = a.perm[p];
        T t = vals[q];
        for (int m = 0; m < a.ndim; ++m) {
          if (m == a.mode) continue;
          const int64_t im = a.coords[(int64_t)m * a.nnz + q];
          t *= ((const T*)a.factors[m])[im * R + lane];
        }
        acc += t;
      }
      ((T*)a.partial)[j * R + lane] = acc;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void mttkrp_coo_rows_kernel(CooArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const T* w = (const T*)a.weights;
  const int R = a.R;
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < a.rows; i += nw) {
    if (lane < R) {
      T acc = (T)0;
      for (int64_t j = a.seg_row_ptr[i]; j < a.seg_row_ptr[i + 1]; ++j) acc += ((const T*)a.partial)[j * R + lane];
      if (w) acc *= w[lane];
      ((T*)a.out)[i * R + lane] = acc;
    }
  }
}

}  // namespace tla

using namespace tla;

extern "C" int tla_khatri_rao(int dtype, const void* const* matrices, int n, const int64_t* rows, int rank,
                              const void* weights, const void* mask, void* out, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "khatri_rao: bad dtype %d", dtype);
  TLA_REQUIRE(n >= 1 && n <= kMaxOrder, "khatri_rao: %d matrices (max %d)", n, kMaxOrder);
  TLA_REQUIRE(rank >= 1, "khatri_rao: rank %d", rank);
  KrArgs a{};
  a.n = n; a.R = rank; a.total = 1;
  for (int m = 0; m < n; ++m) {
    TLA_REQUIRE(rows[m] >= 0 && matrices[m] != nullptr, "khatri_rao: matrix %d", m);
    a.rows[m] = rows[m];
    a.mats[m] = matrices[m];
    a.total *= rows[m];
  }
  a.weights = weights; a.mask = mask;
  const int64_t ne = a.total * rank;
  if (ne == 0) return TLA_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TLA_F32) khatri_rao_kernel<float><<<ceil_div(ne, 256), 256, 0, st>>>(a, (float*)out);
  else khatri_rao_kernel<double><<<ceil_div(ne, 256), 256, 0, st>>>(a, (double*)out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" int tla_mode_dot(int dtype, const void* X, int ndim, const int64_t* shape, const void* U, int64_t J,
                            int is_vector, int mode, int transpose, int precision, void* out, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "mode_dot: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 1 && ndim <= kMaxOrder, "mode_dot: order %d", ndim);
  TLA_REQUIRE(mode >= 0 && mode < ndim, "mode_dot: mode %d", mode);
  TLA_REQUIRE(X && U && out, "mode_dot: NULL pointer");
  int64_t Pa = 1, Pb = 1;
  for (int m = 0; m < mode; ++m) Pa *= shape[m];
  for (int m = mode + 1; m < ndim; ++m) Pb *= shape[m];
  const int64_t Im = shape[mode];
  if (is_vector) J = 1;
  TLA_REQUIRE(J >= 1, "mode_dot: J = %lld", (long long)J);
  // U'(j, m): matrix J x Im (transpose=0: U row-major J x Im; transpose=1:
  // U row-major Im x J); vector: U[m].
  int64_t ldj, ldm;
  if (is_vector) { ldj = 0; ldm = 1; }
  else if (transpose) { ldj = 1; ldm = J; }
  else { ldj = Im; ldm = 1; }
  hipStream_t st = (hipStream_t)stream;
  const int64_t ne = Pa * J * Pb;
  if (ne == 0) return TLA_OK;
  const bool mfma = dtype == TLA_F32 && !is_vector && J <= 64 &&
                    (precision == TLA_PREC_FP32 || precision == TLA_PREC_BF16 || precision == TLA_PREC_BF16X3);
  if (mfma && Pb == 1 && Im % 32 == 0) {
    TtmRowsArgs a{(const float*)X, Pa, Im, (const float*)U, ldj, ldm, (int)J, (float*)out,
                  (int)std::min<int64_t>(Im, 128)};
    if (precision == TLA_PREC_FP32) return launch_ttm_rows<TLA_PREC_FP32>(a, st);
    if (precision == TLA_PREC_BF16) return launch_ttm_rows<TLA_PREC_BF16>(a, st);
    return launch_ttm_rows<TLA_PREC_BF16X3>(a, st);
  }
  if (mfma && Pb > 1 && Pb % 4 == 0 && Pa <= 65535) {
    TtmColsArgs a{(const float*)X, Pa, Im, Pb, (const float*)U, ldj, ldm, (int)J, (float*)out};
    if (precision == TLA_PREC_FP32) return launch_ttm_cols<TLA_PREC_FP32>(a, st);
    if (precision == TLA_PREC_BF16) return launch_ttm_cols<TLA_PREC_BF16>(a, st);
    return launch_ttm_cols<TLA_PREC_BF16X3>(a, st);
  }
  if (dtype == TLA_F32)
    ttm_generic_kernel<float><<<ceil_div(ne, 256), 256, 0, st>>>((const float*)X, Pa, Im, Pb, (const float*)U, J,
                                                                 ldj, ldm, (float*)out);
  else
    ttm_generic_kernel<double><<<ceil_div(ne, 256), 256, 0, st>>>((const double*)X, Pa, Im, Pb, (const double*)U,
                                                                  J, ldj, ldm, (double*)out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" size_t tla_rankdiag_contract_workspace_size(int64_t I, int64_t J, int rank, int axis) {
  if (axis != 0 || rank % 4 != 0) return 0;
  const int64_t S = axis0_splits(I, J, rank);
  return S > 1 ? align256(size_t(S) * J * rank * 4) : 0;
}

extern "C" int tla_rankdiag_contract(const float* Y, int64_t I, int64_t J, int rank, const float* F, int axis,
                                     float* out, void* ws, size_t ws_bytes, void* stream) {
  TLA_REQUIRE(rank >= 1 && rank <= 256, "rankdiag: rank %d", rank);
  TLA_REQUIRE(axis == 0 || axis == 1, "rankdiag: axis %d", axis);
  TLA_REQUIRE(ws_bytes >= tla_rankdiag_contract_workspace_size(I, J, rank, axis), "rankdiag: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  const bool v4 = rank % 4 == 0 && rank <= 64 && (((uintptr_t)Y | (uintptr_t)F | (uintptr_t)out) & 15) == 0;
  if (axis == 1) {
    TLA_REQUIRE(I <= 0x7fffffff, "rankdiag: I too large");
    if (v4)
      rankdiag_axis1_v4_kernel<<<(unsigned)I, 256, 0, st>>>(Y, J, rank, F, out);
    else
      rankdiag_axis1_kernel<<<(unsigned)I, 256, 0, st>>>(Y, J, rank, F, out);
    TLA_LAUNCH_CHECK();
    return TLA_OK;
  }
  const int64_t n = J * rank;
  const int64_t S = v4 ? axis0_splits(I, J, rank) : 1;
  if (v4 && S > 1 && ws && (((uintptr_t)ws) & 15) == 0) {
    const int64_t per = (I + S - 1) / S;
    const dim3 grid((unsigned)ceil_div(n / 4, 64), (unsigned)S);
    rankdiag_axis0_v4_kernel<<<grid, 256, 0, st>>>(Y, I, J, rank, per, F, (float*)ws);
    TLA_LAUNCH_CHECK();
    strided_reduce((const float*)ws, 1, n, S, 0, n, out, st);
    TLA_LAUNCH_CHECK();
    return TLA_OK;
  }
  rankdiag_axis0_kernel<<<(unsigned)ceil_div(n, 64), 256, 0, st>>>(Y, I, J, rank, F, out);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}

extern "C" size_t tla_mttkrp_coo_workspace_size(int dtype, int64_t n_seg, int rank) {
  return (size_t)(n_seg > 0 ? n_seg : 1) * (size_t)rank * (dtype == TLA_F64 ? 8 : 4);
}

extern "C" int tla_mttkrp_coo(int dtype, int64_t nnz, int ndim, const int64_t* shape, const int64_t* coords,
                              const void* values, const int64_t* perm, const int64_t* row_ptr, int64_t n_seg,
                              const int64_t* seg_row, const int64_t* seg_start, const int64_t* seg_row_ptr,
                              int64_t seg_len, const void* const* factors, const void* weights, int rank, int mode,
                              void* out, void* workspace, size_t workspace_bytes, void* stream) {
  TLA_REQUIRE(dtype == TLA_F32 || dtype == TLA_F64, "mttkrp_coo: bad dtype %d", dtype);
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "mttkrp_coo: order %d", ndim);
  TLA_REQUIRE(mode >= 0 && mode < ndim, "mttkrp_coo: mode %d", mode);
  TLA_REQUIRE(rank >= 1 && rank <= 64, "mttkrp_coo: rank %d not in [1, 64]", rank);
  TLA_REQUIRE(nnz >= 0 && n_seg >= 0 && seg_len >= 1, "mttkrp_coo: nnz / segments");
  TLA_REQUIRE(shape && out && row_ptr && seg_row_ptr, "mttkrp_coo: NULL pointer");
  TLA_REQUIRE(n_seg == 0 || (coords && values && perm && seg_row && seg_start && workspace),
              "mttkrp_coo: NULL nonzero / segment arrays");
  TLA_REQUIRE(workspace_bytes >= tla_mttkrp_coo_workspace_size(dtype, n_seg, rank), "mttkrp_coo: workspace too small");
  CooArgs a{};
  for (int m = 0; m < ndim; ++m) {
    TLA_REQUIRE(shape[m] >= 1, "mttkrp_coo: shape[%d] = %lld", m, (long long)shape[m]);
    TLA_REQUIRE(m == mode || factors[m] != nullptr, "mttkrp_coo: factor %d is NULL", m);
    a.factors[m] = factors[m];
  }
  a.coords = coords; a.nnz = nnz; a.vals = values; a.perm = perm; a.row_ptr = row_ptr;
  a.seg_row = seg_row; a.seg_start = seg_start; a.seg_row_ptr = seg_row_ptr; a.n_seg = n_seg; a.seg_len = seg_len;
  // order 2: the Khatri-Rao product of the one other factor carries no
  // weights (tensorly/tenalg/core_tenalg/_khatri_rao.py:67-69)
  a.weights = ndim == 2 ? nullptr : weights; a.ndim = ndim; a.mode = mode; a.R = rank; a.rows = shape[mode];
  a.partial = workspace; a.out = out;
  hipStream_t st = (hipStream_t)stream;
  if (n_seg > 0) {
    const int64_t blocks = std::min<int64_t>((n_seg + 3) / 4, 8192);
    if (dtype == TLA_F32) mttkrp_coo_segments_kernel<float><<<(unsigned)blocks, 256, 0, st>>>(a);
    else mttkrp_coo_segments_kernel<double><<<(unsigned)blocks, 256, 0, st>>>(a);
    TLA_LAUNCH_CHECK();
  }
  const int64_t rblocks = std::min<int64_t>((a.rows + 3) / 4, 8192);
  if (dtype == TLA_F32) mttkrp_coo_rows_kernel<float><<<(unsigned)rblocks, 256, 0, st>>>(a);
  else mttkrp_coo_rows_kernel<double><<<(unsigned)rblocks, 256, 0, st>>>(a);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
