// Error reporting shared by every entry point of the library.
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace tla {

__global__ void __launch_bounds__(256) strided_reduce_kernel(const float* __restrict__ in, int64_t E, int64_t K,
                                                             int64_t gs, int64_t ks, float* __restrict__ out) {
  __shared__ float red[4][64];
  const int el = threadIdx.x & 63, kl = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + el;
  const int64_t g = blockIdx.y;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (e < E) {
    const float* base = in + g * gs + e;
    int64_t k = kl;
    for (; k + 12 < K; k += 16) {
      s0 += base[k * ks];
      s1 += base[(k + 4) * ks];
      s2 += base[(k + 8) * ks];
      s3 += base[(k + 12) * ks];
    }
    for (; k < K; k += 4) s0 += base[k * ks];
  }
  red[kl][el] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (kl == 0 && e < E) out[g * E + e] = (red[0][el] + red[1][el]) + (red[2][el] + red[3][el]);
}

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}
}  // namespace tla

extern "C" const char* tla_last_error(void) { return tla::g_last_error; }

extern "C" int tla_abi_version(void) { return 1; }
