// Allocation-free MTTKRP: tla_mttkrp with its scratch space taken from a
// fixed device array instead of a caller workspace, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// `scratch` is a file-scope device array; every GPU the library runs on holds
// its own copy for the life of the process.  Its device address comes from
// hipGetSymbolAddress, called once per device with that device current (the
// array's name is only a host-side stand-in and never reaches a kernel).
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (different streams, or different host threads) race on it.  The MTTKRP
// kernels write every scratch word they later read, so nothing has to be
// cleared between calls.
#include <mutex>

#include "common.h"

namespace tla {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of this device's `scratch` (looked up once per device).
static int scratch_for_current_device(void** out) {
  int dev = -1;
  TLA_HIP(hipGetDevice(&dev));
  TLA_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "mttkrp_noalloc: device %d beyond %d", dev, kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TLA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TLA_REQUIRE(p != nullptr, "mttkrp_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TLA_OK;
}

}  // namespace tla

using namespace tla;

extern "C" size_t tla_mttkrp_noalloc_max_scratch(void) { return kNoAllocScratchBytes; }

extern "C" int tla_mttkrp_noalloc(int x_dtype, const void* X, int ndim, const int64_t* shape,
                                  const void* const* factors, const void* weights, int rank, int mode, int precision,
                                  void* out, void* stream) {
  TLA_REQUIRE(ndim >= 1 && ndim <= kMaxOrder && shape, "mttkrp_noalloc: bad order %d", ndim);
  const int prec = x_dtype == TLA_F64 ? TLA_PREC_VALU : precision;
  const size_t need = tla_mttkrp_workspace_size(x_dtype, ndim, shape, rank, mode, prec);
  TLA_REQUIRE(need <= kNoAllocScratchBytes,
              "mttkrp_noalloc: this MTTKRP needs %zu bytes of scratch; the allocation-free entry point holds "
              "%zu (use tla_mttkrp with a workspace)",
              need, kNoAllocScratchBytes);
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != TLA_OK) return rc;
  return tla_mttkrp(x_dtype, X, ndim, shape, factors, weights, rank, mode, prec, out, ws, kNoAllocScratchBytes,
                    stream);
}
