// Shared helpers for the gfx950 kernels of the CP-ALS / Tucker-HOOI path.
//
// Conventions used by every kernel in this library:
//  * dense tensors are C-order (last mode contiguous), factors are row-major
//    (I_m x R) -- the layouts tensorly's numpy/pytorch backends hand over;
//  * every entry point returns an int status (0 = ok) and never allocates:
//    scratch memory ("workspace") is passed in by the caller, sized by the
//    matching *_workspace_size() query;
//  * reductions are two-stage with a fixed order, so repeated calls are
//    bitwise identical (no floating-point atomics anywhere).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/tensorly_amd.h"

namespace tla {

void set_error(const char* fmt, ...);

#define TLA_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      ::tla::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,                  \
                       hipGetErrorString(_e));                                     \
      return TLA_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

#define TLA_LAUNCH_CHECK() TLA_HIP(hipGetLastError())

#define TLA_REQUIRE(cond, ...)                                                     \
  do {                                                                             \
    if (!(cond)) {                                                                 \
      ::tla::set_error(__VA_ARGS__);                                               \
      return TLA_ERR_ARG;                                                          \
    }                                                                              \
  } while (0)

constexpr int kMaxOrder = 8;
constexpr int kMaxRank = 64;  // CP rank limit of the ALS solve / Gram kernels
constexpr int kWave = 64;

// A side job for a persistent MTTKRP launch: invert the normal-equations
// matrix of `mode` (see inverse.h) in one extra workgroup.  R == 0: none.
struct InvJob {
  const double* grams;  // ndim Grams, R x R FP64, back to back
  const void* weights;  // nullable; dtype wdt (TLA_F32 / TLA_F64)
  int wdt;
  double l2;
  int ndim, mode, R;
  double* vi;  // R x R FP64 out
  int* vist;   // status out
};

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma_f32_16x16x4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma_bf16_16x16x32(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Round-to-nearest-even split x = hi + lo (+ O(2^-16 |x|)).
__device__ __forceinline__ void split_bf16(const floatx8& x, bf16x8& hi, bf16x8& lo) {
  hi = __builtin_convertvector(x, bf16x8);
  floatx8 back = __builtin_convertvector(hi, floatx8);
  lo = __builtin_convertvector(x - back, bf16x8);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// out[g * E + e] = sum_{k < K} in[g * gs + k * ks + e]  for e < E, g < gridDim.y.
// Block = 4 k-lanes x 64 elements; each lane keeps four independent partial
// sums (so loads overlap), combined in a fixed order: deterministic.
__global__ void __launch_bounds__(256) strided_reduce_kernel(const float* __restrict__ in, int64_t E, int64_t K,
                                                             int64_t gs, int64_t ks, float* __restrict__ out);

inline void strided_reduce(const float* in, int64_t groups, int64_t E, int64_t K, int64_t gs, int64_t ks, float* out,
                           hipStream_t st) {
  for (int64_t g0 = 0; g0 < groups; g0 += 65535) {  // grid.y holds at most 65535 groups
    const int64_t ng = groups - g0 < 65535 ? groups - g0 : 65535;
    dim3 grid((unsigned)ceil_div(E, 64), (unsigned)ng);
    strided_reduce_kernel<<<grid, 256, 0, st>>>(in + g0 * gs, E, K, gs, ks, out + g0 * E);
  }
}

// Chunks C (a power of two dividing K, with K / C >= 16 terms each) that
// strided_reduce_wide splits a sum of K terms into, so that its first pass
// has about 1024 workgroups.  1 when the groups alone give that many, or
// when the groups do not follow one another in memory (gs != K ks).
inline int64_t reduce_chunks(int64_t groups, int64_t E, int64_t K, int64_t gs, int64_t ks) {
  if (groups > 1 && gs != K * ks) return 1;
  const int64_t blocks = groups * ((E + 63) / 64);
  int64_t C = 1;
  while (blocks * C < 1024 && K % (2 * C) == 0 && K / (2 * C) >= 16) C *= 2;
  return C;
}

// floats of temporary storage strided_reduce_wide needs
inline size_t reduce_tmp_elems(int64_t groups, int64_t E, int64_t K, int64_t gs, int64_t ks) {
  const int64_t C = reduce_chunks(groups, E, K, gs, ks);
  return C > 1 ? size_t(groups * C * E) : 0;
}

// strided_reduce in two passes when the groups alone are too few to fill
// the GPU: partial sums over C chunks of the K terms into tmp, then the sum
// of the C chunk sums.  The order of the additions is fixed.
inline void strided_reduce_wide(const float* in, int64_t groups, int64_t E, int64_t K, int64_t gs, int64_t ks,
                                float* out, float* tmp, hipStream_t st) {
  const int64_t C = reduce_chunks(groups, E, K, gs, ks);
  if (C == 1) {
    strided_reduce(in, groups, E, K, gs, ks, out, st);
    return;
  }
  const int64_t Kc = K / C;
  strided_reduce(in, groups * C, E, Kc, Kc * ks, ks, tmp, st);
  strided_reduce(tmp, groups, E, C, C * E, E, out, st);
}

// tla_cp_solve with the driver's options: the relative error of the
// updated model written to errors[err_slot] by the finalising block (last
// mode of a sweep); V^-1 taken from vi_pre when *vi_status >= 0; and, with
// next_mode >= 0, V^-1 of that mode (Grams including this update) written
// to vi_next / vist_next (status 0 SPD elimination, 1 pivoting fallback, 2
// singular) by the finalising block.
int cp_solve_launch(int dtype, const double* const* grams, int ndim, int mode, const void* weights, double l2_reg,
                    const void* mttkrp, int64_t rows, int rank, void* factor_out, double* gram_out, double* iprod_out,
                    int* status_out, int* counter, void* ws, size_t ws_bytes, const double* normx2, double* errors,
                    int err_slot, const double* vi_pre, const int* vi_status, int next_mode, double* vi_next,
                    int* vist_next, int* status_max, void* stream);

// tla_mttkrp (ke == kb + 1) / tla_mttkrp_partial keeping modes [kb, ke),
// with an optional side job (InvJob, nullable) that the persistent kernels
// run in one extra workgroup; *job_taken (nullable) reports whether this
// launch ran it (otherwise the caller must do it some other way).
int mttkrp_launch(int x_dtype, const void* X, int ndim, const int64_t* shape, const void* const* factors,
                  const void* weights, int rank, int kb, int ke, int precision, void* out, void* ws, size_t ws_bytes,
                  const InvJob* job, bool* job_taken, hipStream_t st);

struct ModeShape {
  int ndim;
  int64_t shape[kMaxOrder];
};

}  // namespace tla
