// Native CP-ALS sweep driver: the mode loop of
// tensorly/decomposition/_cp.py:394-531 issued from C++ so that one host call
// runs any number of sweeps with no host synchronisation.
//
// Per sweep and per free mode n:
//   M = MTTKRP(X, factors, weights, n)          (fused kernel, one tensor pass)
//   factors[n] = M V^-1, V = (*_{m != n} G_m + l2 I) * w w^T   (FP64 SPD solve)
//   G_n = factors[n]^T factors[n]                (FP64, fused into the solve)
// and after the last mode (whose MTTKRP is kept) the error of
// _cp.py:217-225 from the Grams: sqrt(|‖X‖² + ‖M‖² − 2<X, M>|) / ‖X‖.
//
// Dimension tree (orders 3 and 4): the first two modes share
// T = X contracted with the factors of modes >= 2 (one pass), the remaining
// modes share the contraction of X with the *updated* first two factors (one
// more pass), so a sweep reads the tensor twice instead of N times.
#include <math.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "extrap.h"

namespace tla {

struct GramPtrs {
  const double* p[kMaxOrder];
};

__global__ void normalize_scales_kernel(GramPtrs grams, int ndim, int R, void* weights,
                                        double* scales /* ndim x R */) {
  // weights fold into factor 0 (tensorly/cp_tensor.py:278-321); the column
  // norms come from the FP64 Gram diagonals.
  const int r = threadIdx.x;
  if (r >= R) return;
  float* w = (float*)weights;
  const double w0 = (double)w[r];
  double neww = 1.0;
  for (int m = 0; m < ndim; ++m) {
    double s = sqrt(fmax(grams.p[m][r * R + r], 0.0));
    if (m == 0) s *= fabs(w0);
    neww *= s;
    // factor m column r is multiplied by f: for m == 0 that is w0 / s.
    const double f = (s == 0.0) ? ((m == 0) ? w0 : 1.0) : ((m == 0) ? w0 / s : 1.0 / s);
    scales[m * R + r] = f;
  }
  w[r] = (float)neww;
}

__global__ void zero_ints_kernel(int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void scale_columns_kernel(float* A, int64_t rows, int R, const double* f) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * R) return;
  A[e] = (float)((double)A[e] * f[e % R]);
}

struct CpAls {
  int x_dtype, ndim, R, precision, dimtree;
  const void* X;
  int64_t shape[kMaxOrder];
  void* factors[kMaxOrder];
  void* weights;
  double l2;
  bool fixed[kMaxOrder];
  // workspace carve-up
  double* grams[kMaxOrder];
  double* normx2;
  double* iprod;
  double* scales;
  int* status;
  int* counter;
  int* status_max;
  double* err_out;  // errors[] of the current sweep (or NULL): fused into the last solve
  int err_slot;
  // V^-1 of the next mode to solve, formed by the finalising block of the
  // previous solve; vi_valid[m]: vi[m] was written since the Grams last
  // changed outside a solve (host-side bookkeeping, stream-ordered)
  double* vi[kMaxOrder];
  int* vist;
  bool vi_valid[kMaxOrder];
  float* M;
  float* T1;
  float* T2;
  void* ws;  // shared scratch for the kernels' own workspaces
  size_t ws_bytes;
  bool initialised;
};

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Layout {
  size_t grams, normx2, iprod, scales, status, vi, vist, M, T1, T2, ws, total;
  size_t ws_bytes;
};

static bool dimtree_ok(int x_dtype, int ndim, const int64_t* shape, int R, int precision) {
  if (!(ndim == 3 || ndim == 4)) return false;
  if (x_dtype == TLA_F64 || precision == TLA_PREC_VALU) return false;
  if (!tla_mttkrp_partial_supported(x_dtype, ndim, shape, R, 0, 2, precision)) return false;
  if (ndim == 4 && !tla_mttkrp_partial_supported(x_dtype, ndim, shape, R, 2, 4, precision)) return false;
  return true;
}

static Layout layout(int x_dtype, int ndim, const int64_t* shape, int R, int precision, int dimtree) {
  Layout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const size_t esz = (x_dtype == TLA_F64) ? 8 : 4;
  L.grams = take(size_t(ndim) * R * R * 8);
  L.normx2 = take(8);
  L.iprod = take(8);
  L.scales = take(size_t(ndim) * R * 8);
  L.status = take(12);
  L.vi = take(size_t(ndim) * R * R * 8);
  L.vist = take(size_t(ndim) * 4);
  int64_t maxI = 0;
  int64_t total = 1;
  for (int m = 0; m < ndim; ++m) { maxI = std::max(maxI, shape[m]); total *= shape[m]; }
  L.M = take(size_t(maxI) * R * esz);
  size_t ws = tla_sqnorm_workspace_size(total);
  for (int m = 0; m < ndim; ++m) {
    ws = std::max(ws, tla_mttkrp_workspace_size(x_dtype, ndim, shape, R, m, precision));
    ws = std::max(ws, tla_cp_solve_workspace_size(shape[m], R));
    ws = std::max(ws, tla_gram64_workspace_size(shape[m], R));
  }
  if (dimtree) {
    L.T1 = take(size_t(shape[0]) * shape[1] * R * 4);
    ws = std::max(ws, tla_mttkrp_partial_workspace_size(x_dtype, ndim, shape, R, 0, 2, precision));
    ws = std::max(ws, tla_rankdiag_contract_workspace_size(shape[0], shape[1], R, 0));
    if (ndim == 4) {
      L.T2 = take(size_t(shape[2]) * shape[3] * R * 4);
      ws = std::max(ws, tla_mttkrp_partial_workspace_size(x_dtype, ndim, shape, R, 2, 4, precision));
      ws = std::max(ws, tla_rankdiag_contract_workspace_size(shape[2], shape[3], R, 0));
    }
  }
  L.ws = take(ws);
  L.ws_bytes = ws;
  L.total = off;
  return L;
}

static void invalidate_inverses(CpAls* h) {
  for (int m = 0; m < kMaxOrder; ++m) h->vi_valid[m] = false;
}

// Where each solve's V^-1 comes from: the MTTKRP launch that precedes the
// solve runs it in one extra workgroup (side job) when its kernel is a
// persistent one; for the second mode of a dimension-tree half (only a short
// rank-diagonal contraction in between) the previous solve's finalising
// block forms it; otherwise the solve inverts V itself.
static InvJob inv_job(const CpAls* h, int n) {
  InvJob j{};
  j.grams = h->grams[0];  // the Grams are back to back in the workspace
  j.weights = h->weights;
  j.wdt = h->x_dtype == TLA_F64 ? TLA_F64 : TLA_F32;
  j.l2 = h->l2;
  j.ndim = h->ndim;
  j.mode = n;
  j.R = h->R;
  j.vi = h->vi[n];
  j.vist = h->vist + n;
  return j;
}

// MTTKRP keeping modes [kb, ke) into out, with the side job for mode
// job_mode (-1: none).
static int mttkrp_with_job(CpAls* h, int kb, int ke, void* out, int job_mode, hipStream_t st) {
  const void* fp[kMaxOrder];
  for (int m = 0; m < h->ndim; ++m) fp[m] = h->factors[m];
  const InvJob job = job_mode >= 0 ? inv_job(h, job_mode) : InvJob{};
  bool taken = false;
  const int rc = mttkrp_launch(h->x_dtype, h->X, h->ndim, h->shape, fp, h->weights, h->R, kb, ke, h->precision, out,
                               h->ws, h->ws_bytes, job_mode >= 0 ? &job : nullptr, &taken, st);
  if (rc == TLA_OK && taken) h->vi_valid[job_mode] = true;
  return rc;
}

// Solve mode n; next >= 0: the finalising block also forms V^-1 of mode next.
static int solve_mode(CpAls* h, int n, const float* M, bool want_iprod, int next, hipStream_t st) {
  const double* gp[kMaxOrder];
  for (int m = 0; m < h->ndim; ++m) gp[m] = h->grams[m];
  const int dt = h->x_dtype == TLA_F64 ? TLA_F64 : TLA_F32;
  const bool pre = h->vi_valid[n];
  invalidate_inverses(h);  // this update changes G_n
  const int rc = cp_solve_launch(dt, gp, h->ndim, n, h->weights, h->l2, M, h->shape[n], h->R, h->factors[n],
                                 h->grams[n], want_iprod ? h->iprod : nullptr, h->status, h->counter, h->ws,
                                 h->ws_bytes, h->normx2, want_iprod ? h->err_out : nullptr, h->err_slot,
                                 pre ? h->vi[n] : nullptr, pre ? h->vist + n : nullptr, next,
                                 next >= 0 ? h->vi[next] : nullptr, next >= 0 ? h->vist + next : nullptr,
                                 h->status_max, st);
  if (rc == TLA_OK && next >= 0) h->vi_valid[next] = true;
  return rc;
}

static int sweep_standard(CpAls* h, hipStream_t st) {
  const int N = h->ndim;
  for (int n = 0; n < N; ++n) {
    if (h->fixed[n]) continue;
    int rc = mttkrp_with_job(h, n, n + 1, h->M, n, st);
    if (rc) return rc;
    rc = solve_mode(h, n, h->M, n == N - 1, -1, st);
    if (rc) return rc;
  }
  return TLA_OK;
}

static int sweep_dimtree(CpAls* h, hipStream_t st) {
  const int N = h->ndim;
  const int R = h->R;
  // First half: modes 0 and 1 share T1 = X x_{m>=2} A_m.
  if (!(h->fixed[0] && h->fixed[1])) {
    int rc = mttkrp_with_job(h, 0, 2, h->T1, h->fixed[0] ? -1 : 0, st);
    if (rc) return rc;
    if (!h->fixed[0]) {
      rc = tla_rankdiag_contract(h->T1, h->shape[0], h->shape[1], R, (const float*)h->factors[1], 1, h->M,
                                 h->ws, h->ws_bytes, st);
      if (rc) return rc;
      rc = solve_mode(h, 0, h->M, false, h->fixed[1] ? -1 : 1, st);
      if (rc) return rc;
    }
    if (!h->fixed[1]) {
      rc = tla_rankdiag_contract(h->T1, h->shape[0], h->shape[1], R, (const float*)h->factors[0], 0, h->M,
                                 h->ws, h->ws_bytes, st);
      if (rc) return rc;
      rc = solve_mode(h, 1, h->M, false, -1, st);
      if (rc) return rc;
    }
  }
  if (N == 3) {
    int rc = mttkrp_with_job(h, 2, 3, h->M, 2, st);
    if (rc) return rc;
    return solve_mode(h, 2, h->M, true, -1, st);
  }
  // N == 4: modes 2 and 3 share T2 = X x_0 A_0 x_1 A_1 (updated factors).
  int rc = mttkrp_with_job(h, 2, 4, h->T2, h->fixed[2] ? 3 : 2, st);
  if (rc) return rc;
  if (!h->fixed[2]) {
    rc = tla_rankdiag_contract(h->T2, h->shape[2], h->shape[3], R, (const float*)h->factors[3], 1, h->M, h->ws,
                               h->ws_bytes, st);
    if (rc) return rc;
    rc = solve_mode(h, 2, h->M, false, 3, st);
    if (rc) return rc;
  }
  rc = tla_rankdiag_contract(h->T2, h->shape[2], h->shape[3], R, (const float*)h->factors[2], 0, h->M, h->ws,
                             h->ws_bytes, st);
  if (rc) return rc;
  return solve_mode(h, 3, h->M, true, -1, st);
}

}  // namespace tla

using namespace tla;

extern "C" size_t tla_cp_als_workspace_size(int x_dtype, int ndim, const int64_t* shape, int rank, int precision,
                                            int dimtree) {
  if (ndim < 2 || ndim > kMaxOrder || rank < 1 || rank > 64) return 0;
  if (x_dtype == TLA_F64) precision = TLA_PREC_VALU;
  if (dimtree && !dimtree_ok(x_dtype, ndim, shape, rank, precision)) dimtree = 0;
  return layout(x_dtype, ndim, shape, rank, precision, dimtree).total;
}

extern "C" int tla_cp_als_dimtree_supported(int x_dtype, int ndim, const int64_t* shape, int rank,
                                            int precision) {
  return dimtree_ok(x_dtype, ndim, shape, rank, precision) ? 1 : 0;
}

extern "C" int tla_cp_als_create(int x_dtype, const void* X, int ndim, const int64_t* shape, int rank,
                                 void* const* factors, void* weights, double l2_reg, int precision,
                                 int dimtree, const int* fixed_modes, void* workspace, size_t ws_bytes,
                                 void** handle_out) {
  TLA_REQUIRE(x_dtype == TLA_F32 || x_dtype == TLA_F64 || x_dtype == TLA_BF16, "cp_als: bad dtype");
  TLA_REQUIRE(ndim >= 2 && ndim <= kMaxOrder, "cp_als: order %d", ndim);
  TLA_REQUIRE(rank >= 1 && rank <= 64, "cp_als: rank %d not in [1, 64]", rank);
  // weights == NULL: unit weights that stay unit (no normalisation); the
  // kernels then skip the weight products entirely
  TLA_REQUIRE(X && factors && workspace && handle_out, "cp_als: NULL pointer");
  if (x_dtype == TLA_F64) precision = TLA_PREC_VALU;
  TLA_REQUIRE(x_dtype != TLA_BF16 || precision == TLA_PREC_BF16, "cp_als: bf16 tensor needs bf16 precision");
  // a bf16 copy has no generic (non-MFMA) MTTKRP: the inner kernels of modes
  // 0..N-2 need the contiguous mode % 32 == 0, the last-mode kernel needs
  // mode N-2 % 8 == 0
  TLA_REQUIRE(x_dtype != TLA_BF16 || (shape[ndim - 1] % 32 == 0 && shape[ndim - 2] % 8 == 0),
              "cp_als: bf16 tensor copy needs shape[%d] %% 32 == 0 and shape[%d] %% 8 == 0 (got %lld, %lld)",
              ndim - 1, ndim - 2, (long long)shape[ndim - 1], (long long)shape[ndim - 2]);
  if (dimtree && !dimtree_ok(x_dtype, ndim, shape, rank, precision)) dimtree = 0;
  const Layout L = layout(x_dtype, ndim, shape, rank, precision, dimtree);
  TLA_REQUIRE(ws_bytes >= L.total, "cp_als: workspace too small (%zu < %zu)", ws_bytes, L.total);
  CpAls* h = new CpAls{};
  h->x_dtype = x_dtype; h->ndim = ndim; h->R = rank; h->precision = precision; h->dimtree = dimtree;
  h->X = X; h->weights = weights; h->l2 = l2_reg;
  char* base = (char*)workspace;
  for (int m = 0; m < ndim; ++m) {
    h->shape[m] = shape[m];
    h->factors[m] = factors[m];
    h->fixed[m] = fixed_modes ? fixed_modes[m] != 0 : false;
    h->grams[m] = (double*)(base + L.grams) + size_t(m) * rank * rank;
  }
  h->fixed[ndim - 1] = false;  // the last mode is never fixed (_cp.py:365-369)
  h->normx2 = (double*)(base + L.normx2);
  h->iprod = (double*)(base + L.iprod);
  h->scales = (double*)(base + L.scales);
  h->status = (int*)(base + L.status);
  // the status slot holds three ints: the last solve's status, the solve's
  // block counter and the largest status since the last refresh
  h->counter = h->status + 1;
  h->status_max = h->status + 2;
  for (int m = 0; m < ndim; ++m) h->vi[m] = (double*)(base + L.vi) + size_t(m) * rank * rank;
  h->vist = (int*)(base + L.vist);
  invalidate_inverses(h);
  h->M = (float*)(base + L.M);
  h->T1 = dimtree ? (float*)(base + L.T1) : nullptr;
  h->T2 = (dimtree && ndim == 4) ? (float*)(base + L.T2) : nullptr;
  h->ws = base + L.ws;
  h->ws_bytes = L.ws_bytes;
  h->initialised = false;
  *handle_out = h;
  return TLA_OK;
}

extern "C" int tla_cp_als_dimtree_active(void* handle) { return handle ? ((CpAls*)handle)->dimtree : 0; }

extern "C" void tla_cp_als_destroy(void* handle) {
  CpAls* h = (CpAls*)handle;
  if (!h) return;
  delete h;
}

// (Re)compute ‖X‖² and the Grams of the current factors (after the caller
// changed the factors, e.g. a line-search step).
extern "C" int tla_cp_als_refresh(void* handle, int with_norm, void* stream) {
  CpAls* h = (CpAls*)handle;
  TLA_REQUIRE(h, "cp_als: NULL handle");
  hipStream_t st = (hipStream_t)stream;
  const int dt = h->x_dtype == TLA_F64 ? TLA_F64 : TLA_F32;
  invalidate_inverses(h);
  zero_ints_kernel<<<1, 64, 0, st>>>(h->status, 3);
  TLA_LAUNCH_CHECK();
  if (with_norm) {
    int64_t total = 1;
    for (int m = 0; m < h->ndim; ++m) total *= h->shape[m];
    TLA_REQUIRE(h->x_dtype != TLA_BF16, "cp_als: norm of a bf16 copy is not computed here");
    int rc = tla_sqnorm(dt, h->X, total, h->normx2, h->ws, h->ws_bytes, st);
    if (rc) return rc;
  }
  for (int m = 0; m < h->ndim; ++m) {
    int rc = tla_gram64(dt, h->factors[m], h->shape[m], h->R, h->grams[m], h->ws, h->ws_bytes, st);
    if (rc) return rc;
  }
  h->initialised = true;
  return TLA_OK;
}

extern "C" int tla_cp_als_set_normx2(void* handle, const double* normx2_device, void* stream) {
  CpAls* h = (CpAls*)handle;
  TLA_REQUIRE(h && normx2_device, "cp_als: NULL pointer");
  TLA_HIP(hipMemcpyAsync(h->normx2, normx2_device, 8, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return TLA_OK;
}

// Run n sweeps; errors[first_slot + s] receives the relative error after
// sweep s (if errors != NULL).  normalize: cp_normalize after each sweep.
extern "C" int tla_cp_als_sweeps(void* handle, int n_sweeps, double* errors, int first_slot, int normalize,
                                 void* stream) {
  CpAls* h = (CpAls*)handle;
  TLA_REQUIRE(h, "cp_als: NULL handle");
  TLA_REQUIRE(h->initialised, "cp_als: call tla_cp_als_refresh first");
  hipStream_t st = (hipStream_t)stream;
  const double* gp[kMaxOrder];
  for (int m = 0; m < h->ndim; ++m) gp[m] = h->grams[m];
  const int dt = h->x_dtype == TLA_F64 ? TLA_F64 : TLA_F32;
  for (int s = 0; s < n_sweeps; ++s) {
    h->err_out = errors;  // the last mode's solve writes errors[first_slot + s]
    h->err_slot = first_slot + s;
    int rc = h->dimtree ? sweep_dimtree(h, st) : sweep_standard(h, st);
    h->err_out = nullptr;
    if (rc) return rc;
    if (normalize) {
      TLA_REQUIRE(dt == TLA_F32, "cp_als: normalize_factors supports float32");
      GramPtrs g{};
      for (int m = 0; m < h->ndim; ++m) g.p[m] = h->grams[m];
      TLA_REQUIRE(h->weights, "cp_als: normalize_factors needs a weights vector");
      invalidate_inverses(h);
      normalize_scales_kernel<<<1, 64, 0, st>>>(g, h->ndim, h->R, h->weights, h->scales);
      TLA_LAUNCH_CHECK();
      for (int m = 0; m < h->ndim; ++m) {
        const int64_t ne = h->shape[m] * h->R;
        scale_columns_kernel<<<ceil_div(ne, 256), 256, 0, st>>>((float*)h->factors[m], h->shape[m], h->R,
                                                                h->scales + m * h->R);
        TLA_LAUNCH_CHECK();
        rc = tla_gram64(dt, h->factors[m], h->shape[m], h->R, h->grams[m], h->ws, h->ws_bytes, st);
        if (rc) return rc;
      }
    }
  }
  return TLA_OK;
}

// Copy part of the driver state to a device buffer (stream-ordered):
// 0 grams (ndim x R x R FP64), 1 ‖X‖² (FP64), 2 last <M, A> (FP64),
// 3 last solve status (int), 4 largest solve status since the last refresh
// (int; 2 = a singular system was met).
extern "C" int tla_cp_als_copy_state(void* handle, int which, void* dst, size_t bytes, void* stream) {
  CpAls* h = (CpAls*)handle;
  TLA_REQUIRE(h && dst, "cp_als: NULL pointer");
  const void* src = nullptr;
  size_t avail = 0;
  switch (which) {
    case 0: src = h->grams[0]; avail = size_t(h->ndim) * h->R * h->R * 8; break;
    case 1: src = h->normx2; avail = 8; break;
    case 2: src = h->iprod; avail = 8; break;
    case 3: src = h->status; avail = 4; break;
    case 4: src = h->status_max; avail = 4; break;
    default: TLA_REQUIRE(false, "cp_als: unknown state %d", which);
  }
  TLA_REQUIRE(bytes <= avail, "cp_als: %zu bytes requested, state %d holds %zu", bytes, which, avail);
  TLA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return TLA_OK;
}

// ---------------------------------------------------------------------------
// Extrapolated ALS (learned step).  Every second sweep the factors are moved
// along the last plain sweep's update, A + alpha (A - A_prev), and the next
// sweep runs from there; the step is kept only if that sweep's error is below
// the error before it, otherwise the saved factors are restored (the sweep
// was wasted).  alpha comes from the MLP over 16 features of the error
// history (or a fixed value).  All decisions are taken on the device.
// ---------------------------------------------------------------------------
namespace tla {

struct FactorPtrs {
  float* p[kMaxOrder];
  int64_t n[kMaxOrder];  // elements
};

// One block: features of sweep s, MLP forward, alpha -> alphas[s + 1].
__global__ void extrap_alpha_kernel(const double* errors, float* alphas, const int* accepted, int s,
                                    const float* params, float alpha_fixed, float alpha_max, float* feat_out,
                                    int write_alpha) {
  __shared__ float f[kMlpIn];
  __shared__ float h1[kMlpH];
  __shared__ float part[kMlpH];
  const int t = threadIdx.x;
  if (t == 0) extrap_features(errors, alphas, accepted, s, f);
  __syncthreads();
  if (feat_out && t < kMlpIn) feat_out[(int64_t)s * kMlpIn + t] = f[t];
  if (!write_alpha) return;
  if (!params) {
    if (t == 0) alphas[s + 1] = alpha_fixed;
    return;
  }
  const float a = mlp_alpha_block(params, f, h1, part, alpha_max);
  if (t == 0) alphas[s + 1] = a;
}

constexpr int kExtrapRows = 64;

struct ExtrapArgs {
  FactorPtrs A, Aprev, Asave;
  double* grams[kMaxOrder];
  double* gsave;  // ndim x R x R
  double* gpart;  // ndim x maxblk x R x R
  int* counters;  // ndim, zero on entry, left at zero
  const float* alphas;
  int slot, R, ndim, maxblk;
  int64_t rows[kMaxOrder];
};

// A_save = A; A = A + alpha (A - A_prev) for kExtrapRows rows of factor
// blockIdx.y, with the FP64 Gram of the new factor built on the way (per-block
// partials, the last block of each factor sums them in block order).  Block
// (0, 0) first saves the current Grams so that a rejected step can restore
// them without recomputing.
__global__ __launch_bounds__(256) void extrapolate_gram_kernel(ExtrapArgs a) {
  __shared__ double tile[kExtrapRows * kMaxRank];
  __shared__ int ticket;
  const int m = blockIdx.y, blk = blockIdx.x, tid = threadIdx.x, R = a.R;
  const int64_t rows = a.rows[m];
  const int nblk = (int)((rows + kExtrapRows - 1) / kExtrapRows);
  if (blk >= nblk) return;
  if (m == 0 && blk == 0)
    for (int e = tid; e < a.ndim * R * R; e += blockDim.x) a.gsave[e] = a.grams[e / (R * R)][e % (R * R)];
  const float alpha = a.alphas[a.slot];
  const int64_t r0 = (int64_t)blk * kExtrapRows;
  const int nr = (int)min((int64_t)kExtrapRows, rows - r0);
  float* A = a.A.p[m] + r0 * R;
  const float* P = a.Aprev.p[m] + r0 * R;
  float* S = a.Asave.p[m] + r0 * R;
  for (int e = tid; e < nr * R; e += blockDim.x) {
    const float x = A[e];
    const float y = x + alpha * (x - P[e]);
    S[e] = x;
    A[e] = y;
    tile[e] = (double)y;
  }
  __syncthreads();
  double* part = a.gpart + ((int64_t)m * a.maxblk + blk) * R * R;
  for (int e = tid; e < R * R; e += blockDim.x) {
    const int r = e / R, q = e % R;
    double s0 = 0.0, s1 = 0.0;
    int i = 0;
    for (; i + 1 < nr; i += 2) {
      s0 = fma(tile[i * R + r], tile[i * R + q], s0);
      s1 = fma(tile[(i + 1) * R + r], tile[(i + 1) * R + q], s1);
    }
    if (i < nr) s0 = fma(tile[i * R + r], tile[i * R + q], s0);
    part[e] = s0 + s1;
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) ticket = atomicAdd(a.counters + m, 1);
  __syncthreads();
  if (ticket != nblk - 1) return;
  __threadfence();
  // partials of the other blocks were published before their tickets; this
  // block never read them, so plain loads (eight blocks in flight) see them
  const double* base = a.gpart + (int64_t)m * a.maxblk * R * R;
  for (int e0 = tid; e0 < R * R; e0 += 4 * 256) {
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = 0; b0 < nblk; b0 += 8) {
      double v[8][4];
#pragma unroll
      for (int bb = 0; bb < 8; ++bb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = e0 + j * 256;
          v[bb][j] = (b0 + bb < nblk && e < R * R) ? base[(int64_t)(b0 + bb) * R * R + e] : 0.0;
        }
#pragma unroll
      for (int bb = 0; bb < 8; ++bb)
#pragma unroll
        for (int j = 0; j < 4; ++j) g[j] += v[bb][j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (e0 + j * 256 < R * R) a.grams[m][e0 + j * 256] = g[j];
  }
  if (tid == 0) a.counters[m] = 0;
}

__global__ void copy_factors_kernel(FactorPtrs dst, FactorPtrs src) {
  const int m = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= src.n[m]) return;
  dst.p[m][e] = src.p[m][e];
}

// Keep the extrapolated step iff errors[s + 1] < errors[s]; else restore the
// saved factors and Grams and carry the previous error forward.
struct GramPtrArray {
  double* p[kMaxOrder];
};

// A_prev of the next pair is the factor kept here, written in the same pass.
__global__ void accept_or_revert_kernel(FactorPtrs A, FactorPtrs Asave, FactorPtrs Aprev, GramPtrArray grams,
                                        const double* gsave, int ndim, int R, double* errors, int* accepted, int s,
                                        int* vist) {
  const bool ok = errors[s + 1] < errors[s];
  const int m = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < A.n[m]) {
    const float v = ok ? A.p[m][e] : Asave.p[m][e];
    if (!ok) A.p[m][e] = v;
    Aprev.p[m][e] = v;
  }
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    if (!ok)
      for (int i = threadIdx.x; i < ndim * R * R; i += blockDim.x) grams.p[i / (R * R)][i % (R * R)] = gsave[i];
    if (threadIdx.x == 0) {
      accepted[s + 1] = ok ? 1 : 0;
      if (!ok) errors[s + 1] = errors[s];
    }
    if (!ok && threadIdx.x < ndim) vist[threadIdx.x] = -1;
  }
}

}  // namespace tla

extern "C" int tla_mlp_num_params(void);

extern "C" size_t tla_cp_als_extrap_workspace_size(void* handle) {
  CpAls* h = (CpAls*)handle;
  if (!h) return 0;
  size_t tot = 0;
  int64_t maxblk = 0;
  for (int m = 0; m < h->ndim; ++m) {
    tot += align256(size_t(h->shape[m]) * h->R * 4);
    maxblk = std::max(maxblk, (h->shape[m] + kExtrapRows - 1) / kExtrapRows);
  }
  const size_t rr = size_t(h->R) * h->R * 8;
  return 2 * tot + align256(h->ndim * rr) + align256(h->ndim * maxblk * rr) + align256(size_t(h->ndim) * 4);
}

extern "C" int tla_cp_als_run_extrapolated(void* handle, int n_sweeps, const float* mlp_params, float alpha_fixed,
                                           float alpha_max, int warmup, double* errors, int* accepted, float* alphas,
                                           float* features, void* ws, size_t ws_bytes, void* stream) {
  CpAls* h = (CpAls*)handle;
  TLA_REQUIRE(h && errors && accepted && alphas, "cp_als_extrapolated: NULL pointer");
  TLA_REQUIRE(h->initialised, "cp_als: call tla_cp_als_refresh first");
  TLA_REQUIRE(h->x_dtype != TLA_F64, "cp_als_extrapolated: float32 factors only");
  TLA_REQUIRE(ws_bytes >= tla_cp_als_extrap_workspace_size(handle), "cp_als_extrapolated: workspace too small");
  TLA_REQUIRE(warmup >= 1, "cp_als_extrapolated: warmup must be >= 1");
  hipStream_t st = (hipStream_t)stream;
  const double* gp[kMaxOrder];
  for (int m = 0; m < h->ndim; ++m) gp[m] = h->grams[m];
  FactorPtrs A{}, Aprev{}, Asave{};
  char* base = (char*)ws;
  size_t off = 0;
  int64_t maxn = 0;
  for (int m = 0; m < h->ndim; ++m) {
    const int64_t n = h->shape[m] * h->R;
    maxn = std::max(maxn, n);
    A.p[m] = (float*)h->factors[m];
    A.n[m] = Aprev.n[m] = Asave.n[m] = n;
    Aprev.p[m] = (float*)(base + off);
    off += align256(size_t(n) * 4);
  }
  for (int m = 0; m < h->ndim; ++m) {
    Asave.p[m] = (float*)(base + off);
    off += align256(size_t(A.n[m]) * 4);
  }
  const int R = h->R;
  int64_t maxblk = 0;
  for (int m = 0; m < h->ndim; ++m) maxblk = std::max(maxblk, (h->shape[m] + kExtrapRows - 1) / kExtrapRows);
  ExtrapArgs ea{};
  ea.A = A;
  ea.Aprev = Aprev;
  ea.Asave = Asave;
  GramPtrArray gpa{};
  for (int m = 0; m < h->ndim; ++m) {
    ea.grams[m] = h->grams[m];
    gpa.p[m] = h->grams[m];
    ea.rows[m] = h->shape[m];
  }
  ea.gsave = (double*)(base + off);
  off += align256(size_t(h->ndim) * R * R * 8);
  ea.gpart = (double*)(base + off);
  off += align256(size_t(h->ndim) * maxblk * R * R * 8);
  ea.counters = (int*)(base + off);
  ea.alphas = alphas;
  ea.R = R;
  ea.ndim = h->ndim;
  ea.maxblk = (int)maxblk;
  zero_ints_kernel<<<1, 64, 0, st>>>(ea.counters, h->ndim);
  TLA_LAUNCH_CHECK();
  const dim3 fgrid((unsigned)ceil_div(maxn, 256), (unsigned)h->ndim);
  const dim3 egrid((unsigned)maxblk, (unsigned)h->ndim);
  auto sweep = [&](int slot) -> int {
    h->err_out = errors;
    h->err_slot = slot;
    const int rc = h->dimtree ? sweep_dimtree(h, st) : sweep_standard(h, st);
    h->err_out = nullptr;
    return rc;
  };
  int s = 0;
  // plain warm-up sweeps (flags/alphas for these slots are set by the caller)
  for (; s < std::min(warmup, n_sweeps); ++s) {
    int rc = sweep(s);
    if (rc) return rc;
  }
  // Per pair: save A_prev, plain sweep, step from the network, extrapolate
  // (+ Grams), sweep from there, keep or restore.  The Grams stay current
  // throughout, so no separate Gram pass is needed.
  // A_prev for the first pair; later pairs get it from accept_or_revert
  copy_factors_kernel<<<fgrid, 256, 0, st>>>(Aprev, A);
  TLA_LAUNCH_CHECK();
  while (s < n_sweeps) {
    int rc = sweep(s);
    if (rc) return rc;
    if (s + 1 >= n_sweeps) break;
    extrap_alpha_kernel<<<1, 64, 0, st>>>(errors, alphas, accepted, s, mlp_params, alpha_fixed, alpha_max,
                                           features, 1);
    TLA_LAUNCH_CHECK();
    ea.slot = s + 1;
    extrapolate_gram_kernel<<<egrid, 256, 0, st>>>(ea);
    TLA_LAUNCH_CHECK();
    invalidate_inverses(h);
    if ((rc = sweep(s + 1))) return rc;
    // on a revert the kernel also marks the inverse formed for the next
    // sweep (from the rejected Grams) as stale
    accept_or_revert_kernel<<<fgrid, 256, 0, st>>>(A, Asave, Aprev, gpa, ea.gsave, h->ndim, R, errors, accepted, s,
                                                   h->vist);
    TLA_LAUNCH_CHECK();
    s += 2;
  }
  return TLA_OK;
}

// Features of sweep s (the network's input) written to out[s * 16 ..], so
// training data uses exactly the inference-time code.
extern "C" int tla_extrap_features(const double* errors, const float* alphas, const int* accepted, int s,
                                   float* out, void* stream) {
  TLA_REQUIRE(errors && alphas && accepted && out && s >= 0, "extrap_features: bad argument");
  tla::extrap_alpha_kernel<<<1, 64, 0, (hipStream_t)stream>>>(errors, const_cast<float*>(alphas), accepted, s,
                                                               nullptr, 0.0f, 0.0f, out, 0);
  TLA_LAUNCH_CHECK();
  return TLA_OK;
}
