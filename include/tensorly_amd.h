/*
 * tensorly_amd -- C ABI of the MI355X (gfx950) CP-ALS / Tucker-HOOI kernels.
 *
 * Plain pointers and sizes only.  Every pointer argument named like a tensor
 * is a DEVICE pointer; shape / pointer-list arguments are HOST arrays.
 * `stream` is a hipStream_t (NULL = default stream).  Entry points never
 * allocate device memory: scratch space ("workspace") is provided by the
 * caller, sized by the matching *_workspace_size() query.  Every entry point
 * returns TLA_OK (0) or an error code; tla_last_error() (thread-local) gives
 * the message.
 *
 * Layout contract (the numpy / pytorch backends' own): dense tensors are
 * C-order with the last mode contiguous; factor matrices are row-major
 * (I_m x R); `unfold(X, n)` columns and Khatri-Rao rows run row-major over the
 * remaining modes in increasing order (tensorly/base.py:39-53,
 * tensorly/tenalg/core_tenalg/_khatri_rao.py:94-105).
 *
 * Interfaces replaced are cited as reference file:line (tensorly 0.9.0).
 */
#ifndef TENSORLY_AMD_H
#define TENSORLY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TLA_OK 0
#define TLA_ERR_ARG 1         /* bad shape / argument: tensorly raises ValueError */
#define TLA_ERR_HIP 2         /* HIP runtime error */
#define TLA_ERR_UNSUPPORTED 3 /* combination not implemented */
#define TLA_ERR_NUMERIC 4     /* singular system / non-finite result */

/* element types */
#define TLA_F32 0
#define TLA_F64 1
#define TLA_BF16 2 /* bf16 storage copy of a float32 tensor (MTTKRP only) */

/* arithmetic of the tensor contraction */
#define TLA_PREC_FP32 0   /* FP32 MFMA, FP32 accumulation */
#define TLA_PREC_BF16 1   /* BF16 MFMA: tensor rounded to bf16 (or a bf16 copy) times the
                             factor / matrix operand split hi + lo, 2 MFMAs, FP32 accum. */
#define TLA_PREC_BF16X3 2 /* split hi+lo BF16, 3 MFMAs per product, FP32 accum. */
#define TLA_PREC_VALU 3   /* plain vector ALU in the tensor's own type */

const char* tla_last_error(void);
int tla_abi_version(void);

/* ---------------------------------------------------------------------
 * MTTKRP:  out(I_mode x R) = unfold(X, mode) @ khatri_rao(factors, weights,
 * skip_matrix=mode), with the Khatri-Rao product never materialised.
 * Replaces tensorly/tenalg/core_tenalg/mttkrp.py:9-50
 * (unfolding_dot_khatri_rao) and tensorly/tenalg/einsum_tenalg/mttkrp.py:6-39.
 *   x_dtype  TLA_F32 / TLA_F64 / TLA_BF16; factors, weights and out are F64
 *            when x_dtype is F64, else F32.
 *   factors  host array of ndim device pointers (entry `mode` is ignored).
 *   weights  device pointer of R values, or NULL for ones.  Ignored when
 *            ndim == 2: khatri_rao of the one remaining matrix returns it
 *            unweighted (tensorly/tenalg/core_tenalg/_khatri_rao.py:67-69).
 * ------------------------------------------------------------------- */
size_t tla_mttkrp_workspace_size(int x_dtype, int ndim, const int64_t* shape, int rank,
                                 int mode, int precision);
int tla_mttkrp(int x_dtype, const void* X, int ndim, const int64_t* shape,
               const void* const* factors, const void* weights, int rank, int mode,
               int precision, void* out, void* workspace, size_t workspace_bytes,
               void* stream);

/* ---------------------------------------------------------------------
 * Allocation-free MTTKRP.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * result as tla_mttkrp, without the workspace: the scratch space is a fixed
 * device array of tla_mttkrp_noalloc_max_scratch() bytes (64 MiB) that each
 * GPU holds for the life of the process; an MTTKRP whose
 * tla_mttkrp_workspace_size() exceeds it returns TLA_ERR_ARG.  The call
 * allocates nothing.  Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (other streams or host threads) race.  The
 * stream must belong to the current device.
 * ------------------------------------------------------------------- */
size_t tla_mttkrp_noalloc_max_scratch(void);
int tla_mttkrp_noalloc(int x_dtype, const void* X, int ndim, const int64_t* shape,
                       const void* const* factors, const void* weights, int rank, int mode,
                       int precision, void* out, void* stream);

/* ---------------------------------------------------------------------
 * Partial MTTKRP (dimension tree): contract every mode except the
 * contiguous block [keep_begin, keep_end) -- which must be a prefix
 * (keep_begin == 0) or a suffix (keep_end == ndim) of the modes -- with the
 * corresponding factor columns.  Output is the (prod kept sizes) x R matrix
 * Y[q, r] = sum_p X[..] * w_r * prod_{m not kept} A_m[i_m, r].
 * Used by the CP-ALS dimension-tree sweep; no tensorly counterpart.
 * ------------------------------------------------------------------- */
size_t tla_mttkrp_partial_workspace_size(int x_dtype, int ndim, const int64_t* shape,
                                         int rank, int keep_begin, int keep_end,
                                         int precision);
/* 1 if the partial contraction has a kernel for this shape / precision */
int tla_mttkrp_partial_supported(int x_dtype, int ndim, const int64_t* shape, int rank,
                                 int keep_begin, int keep_end, int precision);
int tla_mttkrp_partial(int x_dtype, const void* X, int ndim, const int64_t* shape,
                       const void* const* factors, const void* weights, int rank,
                       int keep_begin, int keep_end, int precision, void* out,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Contract a rank-diagonal 3-way array Y (I x J x R) with one factor:
 * contract_axis 1: out[i,r] = sum_j Y[i,j,r] F[j,r]
 * contract_axis 0: out[j,r] = sum_i Y[i,j,r] F[i,r]
 * (second level of the dimension tree). F32 only. */
size_t tla_rankdiag_contract_workspace_size(int64_t I, int64_t J, int rank, int contract_axis);
int tla_rankdiag_contract(const float* Y, int64_t I, int64_t J, int rank, const float* F,
                          int contract_axis, float* out, void* workspace,
                          size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * MTTKRP of a sparse tensor in coordinate form (the unfolding_dot_khatri_rao
 * of tensorly/contrib/sparse/cp_tensor.py:9, there over the pydata `sparse`
 * backend): out(I_mode x R)[i, r] = w[r] * sum over nonzeros with index i in
 * `mode` of val * prod_{m != mode} factors[m][i_m, r].
 *   coords   device int64 [ndim][nnz] (mode-major); values nnz in dtype;
 *   perm     device int64 nnz: the nonzeros ordered by their `mode` index
 *            (a stable order makes the result repeatable bit for bit);
 *   row_ptr  device int64 shape[mode] + 1: row i's nonzeros are
 *            perm[row_ptr[i] .. row_ptr[i+1]);
 *   segments each row's nonzeros cut into n_seg pieces of at most seg_len:
 *            seg_row[j] / seg_start[j] (device int64, n_seg) give piece j's
 *            row and first position in perm; seg_row_ptr (shape[mode] + 1)
 *            delimits each row's pieces;
 *   factors  host array of ndim device pointers (entry `mode` unused), dtype;
 *   weights  R values (dtype) or NULL (ignored when ndim == 2, as for
 *            tla_mttkrp).  rank <= 64.
 *   workspace  tla_mttkrp_coo_workspace_size(dtype, n_seg, rank) bytes.
 * ------------------------------------------------------------------- */
size_t tla_mttkrp_coo_workspace_size(int dtype, int64_t n_seg, int rank);
int tla_mttkrp_coo(int dtype, int64_t nnz, int ndim, const int64_t* shape, const int64_t* coords,
                   const void* values, const int64_t* perm, const int64_t* row_ptr, int64_t n_seg,
                   const int64_t* seg_row, const int64_t* seg_start, const int64_t* seg_row_ptr,
                   int64_t seg_len, const void* const* factors, const void* weights, int rank, int mode,
                   void* out, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Khatri-Rao product (materialised, for the khatri_rao API only).
 * Replaces tensorly/tenalg/core_tenalg/_khatri_rao.py:9-109.
 *   matrices host array of n device pointers (rows[i] x R, row-major);
 *   weights  R values or NULL; mask prod(rows) values or NULL.
 * ------------------------------------------------------------------- */
int tla_khatri_rao(int dtype, const void* const* matrices, int n_matrices, const int64_t* rows,
                   int rank, const void* weights, const void* mask, void* out, void* stream);

/* ---------------------------------------------------------------------
 * n-mode product (TTM / TTV).  Replaces
 * tensorly/tenalg/core_tenalg/n_mode_product.py:5-76 (mode_dot).
 *   Matrix case: U is (J x I_mode) (transpose=0) or (I_mode x J)
 *   (transpose=1); out has shape[mode] replaced by J.
 *   Vector case (is_vector=1): U holds I_mode values; out drops the mode.
 * ------------------------------------------------------------------- */
int tla_mode_dot(int dtype, const void* X, int ndim, const int64_t* shape, const void* U,
                 int64_t J, int is_vector, int mode, int transpose, int precision, void* out,
                 void* stream);

/* ---------------------------------------------------------------------
 * FP64 Gram G = A^T A of a factor (rows x R), the A_i^T A_i products of
 * tensorly/decomposition/_cp.py:411-416 and tensorly/cp_tensor.py:635-637.
 * ------------------------------------------------------------------- */
size_t tla_gram64_workspace_size(int64_t rows, int rank);
int tla_gram64(int dtype, const void* A, int64_t rows, int rank, double* G, void* workspace,
               size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * ALS factor update (tensorly/decomposition/_cp.py:411-427):
 *   V = (prod_{i != mode} G_i + l2_reg I) * (w w^T)          (FP64)
 *   factor = mttkrp @ inv(V)          (FP64 Gauss-Jordan elimination
 *                                      without pivoting, valid for SPD V;
 *                                      partial pivoting if a pivot is not
 *                                      positive; products on FP64 MFMA)
 * and, fused, the FP64 Gram of the new factor (gram_out) and, if
 * iprod_out != NULL, <mttkrp, factor> (the error_calc inner product,
 * _cp.py:222).  grams: host array of ndim device pointers to R x R FP64.
 * Returns TLA_ERR_NUMERIC only through `status_out` (device int, nullable):
 * 0 ok, 1 LU fallback used, 2 singular.
 * `counter` (device int, nullable) must hold 0 on entry and is left at 0: the
 * last workgroup then sums the per-block partials itself (one launch instead
 * of two); NULL uses a separate finalising launch.
 * ------------------------------------------------------------------- */
size_t tla_cp_solve_workspace_size(int64_t rows, int rank);
int tla_cp_solve(int dtype, const double* const* grams, int ndim, int mode,
                 const void* weights, double l2_reg, const void* mttkrp, int64_t rows, int rank,
                 void* factor_out, double* gram_out, double* iprod_out, int* status_out,
                 int* counter, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Squared Frobenius norm with FP64 accumulation (the norm of
 * tensorly/backend/core.py:714-739, order 2, before the sqrt).
 * ------------------------------------------------------------------- */
size_t tla_sqnorm_workspace_size(int64_t n);
int tla_sqnorm(int dtype, const void* X, int64_t n, double* out, void* workspace,
               size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * CP reconstruction error from Grams (tensorly/decomposition/_cp.py:217-225
 * with cp_norm, tensorly/cp_tensor.py:614-644):
 *   err = sqrt(|normx2 + sum((prod_m G_m) * w w^T) - 2 * iprod|)
 *   errors[slot] = err / sqrt(normx2)
 * All operands are device FP64 scalars / matrices; weights in `dtype`.
 * ------------------------------------------------------------------- */
int tla_cp_error(const double* const* grams, int ndim, const void* weights, int dtype, int rank,
                 const double* normx2, const double* iprod, double* errors, int slot,
                 void* stream);

/* ---------------------------------------------------------------------
 * Multiplicative update of one factor of a non-negative CP model
 * (tensorly/decomposition/_nn_cp.py:111-135, the mode loop of
 * non_negative_parafac):
 *   accum  = (prod_{m != mode} G_m) * w w^T          (FP64, R x R)
 *   factor = factor * max(mttkrp, eps) / max(factor accum, eps)
 * in place.  factor and mttkrp are rows x rank in `dtype` (TLA_F32 /
 * TLA_F64); grams is a host array of ndim device pointers to FP64 R x R
 * Grams (entry `mode` unused); weights (nullable: ones) in `wdtype`.
 * rank <= 64.
 * ------------------------------------------------------------------- */
int tla_nn_mu_update(int dtype, void* factor, const void* mttkrp, int64_t rows, int rank,
                     const double* const* grams, int ndim, int mode, const void* weights, int wdtype,
                     double eps, void* stream);

/* ---------------------------------------------------------------------
 * Gradient of 0.5 ||X - [[w; A]]||^2 with respect to factor `mode`
 * (tensorly/cp_tensor.py:377-430, cp_lstsq_grad without a mask), from the
 * tensor's MTTKRP and the factors' Grams instead of the residual tensor:
 *   grad = factor accum - mttkrp,  accum = (prod_{m != mode} G_m) * w w^T
 * (-MTTKRP(X - M) = MTTKRP(M) - MTTKRP(X), MTTKRP(M) = factor accum).
 * Arguments as tla_nn_mu_update; grad (rows x rank, dtype) must not alias.
 * ------------------------------------------------------------------- */
int tla_cp_lstsq_grad_mode(int dtype, const void* factor, const void* mttkrp, void* grad, int64_t rows,
                           int rank, const double* const* grams, int ndim, int mode, const void* weights,
                           int wdtype, void* stream);

/* ---------------------------------------------------------------------
 * HALS non-negative least squares for one CP factor
 * (tensorly/solvers/nnls.py:5-173 hals_nnls, as called by
 * non_negative_parafac_hals, tensorly/decomposition/_nn_cp.py:300-311):
 * with UtU = (prod_{m != mode} G_m) * w w^T and UtM = mttkrp^T, the rows k of
 * V = factor^T are updated in order,
 *   V[k,:] = max((UtM[k,:] - UtU[k,:] V + UtU[k,k] V[k,:] - s) / UtU[k,k], 0)
 * (s = sparsity when has_sparsity; rows with UtU[k,k] == 0 are skipped), for
 * at most n_iter_max iterations, stopping when an iteration's error sum is
 * below tol times the first iteration's; each row update adds
 * ||V - newV||^2 with the new row broadcast against every row of the current
 * V, as the reference does (tensorly/solvers/nnls.py:147).  factor (rows x rank, dtype) is
 * updated in place, in FP64 arithmetic; iters_out (device int, nullable)
 * receives the iterations run.  With a workspace (nullable) and
 * n_iter_max <= 1000 each iteration is one launch over the whole GPU (a
 * device-side flag ends the remaining launches early); otherwise one
 * workgroup runs the loop.  Arguments otherwise as tla_nn_mu_update.
 * ------------------------------------------------------------------- */
size_t tla_hals_nnls_workspace_size(int64_t rows, int rank);
int tla_hals_nnls(int dtype, void* factor, const void* mttkrp, int64_t rows, int rank,
                  const double* const* grams, int ndim, int mode, const void* weights, int wdtype,
                  double sparsity, int has_sparsity, int n_iter_max, double tol, int* iters_out,
                  void* workspace, size_t workspace_bytes, void* stream);


/* ---------------------------------------------------------------------
 * Masked imputation of parafac(..., mask=...)
 * (tensorly/decomposition/_cp.py:194-207, :437-445): with the CP model
 * M = [[w; factors]] (never formed), X_new = X * mask + M * (1 - mask);
 * sums[0] = ||X_new||^2 and sums[1] = ||X_new - M||^2 (device FP64); X is
 * overwritten with X_new when `write`.  X and mask are dense C-order in
 * dtype; factors a host array of ndim device pointers; weights (R values)
 * nullable; rank <= 64.  Workspace tla_cp_impute_workspace_size bytes.
 * tla_masked_blend: out = X * mask + low_rank * (1 - mask) elementwise (the
 * SVD mask imputation of tensorly/tenalg/svd.py:431-439).
 * ------------------------------------------------------------------- */
size_t tla_cp_impute_workspace_size(int dtype, int ndim, const int64_t* shape);
int tla_cp_impute(int dtype, void* X, const void* mask, int ndim, const int64_t* shape,
                  const void* const* factors, const void* weights, int rank, int write, double* sums,
                  void* workspace, size_t workspace_bytes, void* stream);
int tla_masked_blend(int dtype, const void* X, const void* mask, const void* low_rank, void* out, int64_t n,
                     void* stream);

/* ---------------------------------------------------------------------
 * Reduced QR of a rows x cols matrix (rows >= cols), the tl.qr that
 * parafac(..., orthogonalise=...) applies to each factor
 * (tensorly/decomposition/_cp.py:395-399; numpy: LAPACK geqrf + orgqr).
 * Householder reflectors with LAPACK's signs (R's diagonal is
 * -sign(alpha) * norm), FP64 arithmetic, one workgroup.  Q (rows x cols,
 * may alias A) and R (cols x cols, upper triangular, nullable) in dtype;
 * workspace tla_qr_workspace_size(rows, cols) bytes.
 * ------------------------------------------------------------------- */
size_t tla_qr_workspace_size(int64_t rows, int cols);
int tla_qr(int dtype, const void* A, int64_t rows, int cols, void* Q, void* R, void* workspace,
           size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Proximal operators of tensorly/tenalg/proximal.py:153-267
 * (proximal_operator), for a rows x cols matrix (cols <= 64), by code:
 * ------------------------------------------------------------------- */
#define TLA_CON_NONE 0
#define TLA_CON_NON_NEGATIVE 1        /* clip(v, 0, max(v)) */
#define TLA_CON_L1 2                  /* soft thresholding by param */
#define TLA_CON_L2 3                  /* block soft thresholding, Frobenius norm */
#define TLA_CON_L2_SQUARE 4           /* v / (1 + 2 param) */
#define TLA_CON_UNIMODALITY 5         /* per column */
#define TLA_CON_NORMALIZE 6           /* v / max|v| */
#define TLA_CON_SIMPLEX 7             /* per column, simplex of radius param */
#define TLA_CON_NORMALIZED_SPARSITY 8 /* hard thresholding, then unit norm */
#define TLA_CON_SOFT_SPARSITY 9       /* per column, l1 ball of radius param */
#define TLA_CON_SMOOTHNESS 10         /* per column tridiagonal solve */
#define TLA_CON_MONOTONICITY 11       /* per column, non-decreasing */
#define TLA_CON_HARD_SPARSITY 12      /* keep the param largest |v| */

/* out = prox(in) in FP64 arithmetic (in / out in dtype, may alias);
 * workspace tla_prox_workspace_size(rows, cols) bytes.  Entries of equal
 * magnitude at the hard-thresholding cut keep the later flat index first. */
size_t tla_prox_workspace_size(int64_t rows, int cols);
int tla_prox(int dtype, const void* in, void* out, int64_t rows, int cols, int constraint, double param,
             void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * ADMM factor update of constrained CP: the inner loop of
 * tensorly/solvers/admm.py:5-155 as called by constrained_parafac
 * (tensorly/decomposition/_constrained_cp.py:357-375), in one launch.
 * With UtU = utu (R x R FP64, device) or, when utu is NULL, the Hadamard
 * product of the other modes' FP64 Grams (host array of ndim device
 * pointers, entry `mode` unused), rho = trace(UtU) / R and UtM = mttkrp
 * (rows x R, dtype), for at most n_iter_max iterations:
 *   x_split^T = (UtM + rho (x + dual)) (UtU + rho I)^-1
 *   x         = prox(x_split^T - dual)                (constraint, param)
 *   dual     += x - x_split^T
 * stopping when ||x - x_split^T|| < tol ||x|| and ||x - x_old|| < tol ||dual||.
 * factor (x, rows x R, dtype) is updated in place; x_split (x_split^T,
 * rows x R FP64) is written; dual (rows x R FP64) is updated.  Device
 * outputs, each nullable: status_out 0, or 2 when UtU + rho I is singular
 * (the reference's tl.solve raises; nothing else is written then);
 * iters_out the iterations run; norms_out[2] = ||x - x_split^T||, ||x|| of
 * the last iteration (the constraint error of _constrained_cp.py:386-389);
 * iprod_out = sum(UtM * x) with x as stored (the error's inner product,
 * _constrained_cp.py:380).  Workspace: tla_admm_workspace_size(rows, R).
 * ------------------------------------------------------------------- */
size_t tla_admm_workspace_size(int64_t rows, int rank);
int tla_admm_update(int dtype, void* factor, double* x_split, double* dual, const void* mttkrp, int64_t rows,
                    int rank, const double* utu, const double* const* grams, int ndim, int mode, int n_iter_max,
                    double tol, int constraint, double param, int* status_out, int* iters_out, double* norms_out,
                    double* iprod_out, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Leading left singular vectors through the FP64 Gram (replaces the LAPACK
 * SVD of tensorly/tenalg/svd.py:211-235 on the Tucker and init='svd' paths,
 * called from tensorly/decomposition/_tucker.py:63-79,194-202 and
 * tensorly/decomposition/_cp.py:72-100).
 *   tla_mode_gram64: G = unfold(X, mode) unfold(X, mode)^T (I_mode^2 FP64),
 *                    without forming the unfolding.
 *   tla_sym_eig:     k leading eigenpairs (descending) of symmetric FP64 G
 *                    (n x n).  n <= 128: Jacobi on the whole matrix (any
 *                    k <= n); n > 128: subspace iteration with Rayleigh-Ritz,
 *                    block 48 / 64 / 128 for k <= 40 / 56 / 120 (larger k
 *                    returns TLA_ERR_UNSUPPORTED), optional warm start Q0
 *                    (n x k0, its first columns used), converged when every
 *                    leading residual is <= tol * lambda_1 (or after
 *                    max_iter steps; the sweep or step count goes to
 *                    *iters_out if non-NULL, device int).
 * ------------------------------------------------------------------- */
size_t tla_mode_gram64_workspace_size(int ndim, const int64_t* shape, int mode);
int tla_mode_gram64(int dtype, const void* X, int ndim, const int64_t* shape, int mode, double* G,
                    void* workspace, size_t workspace_bytes, void* stream);
size_t tla_sym_eig_workspace_size(int n);
int tla_sym_eig(const double* G, int n, int k, const double* Q0, int k0, double* evals,
                double* evecs, double tol, int max_iter, int* iters_out, void* workspace,
                size_t workspace_bytes, void* stream);
/* count (<= 8) independent tla_sym_eig problems in one launch per solver
 * kind, one workgroup each (the three SVDs of the Tucker init,
 * tensorly/decomposition/_tucker.py:63-79, run side by side).  Per-problem
 * arrays; Q0 and k0 may be NULL (no warm starts), entries of Q0 NULL. */
int tla_sym_eig_batch(int count, const double* const* G, const int* n, const int* k,
                      const double* const* Q0, const int* k0, double* const* evals,
                      double* const* evecs, double tol, int max_iter, void* const* workspace,
                      const size_t* workspace_bytes, void* stream);
/* tla_sym_eig with the rest of the reference's truncated-SVD path fused into
 * the same launch: the k leading left singular vectors U (n x k, row-major,
 * dtype TLA_F32 / TLA_F64) of a matrix whose row Gram is G, with the sign
 * rule of svd_flip(u_based_decision=True) (tensorly/tenalg/svd.py:13-42:
 * each column times the sign of its entry of largest magnitude, the first
 * such row on ties), s = sqrt(max(eigenvalue, 0)) (k, dtype; nullable) and
 * U64 (the same vectors in FP64; nullable).  For n > 128,
 * tla_gram_svd_block(n, k) = b is the subspace block: W_out (n x b FP64,
 * nullable) receives the final block, and W0 = that block with k0 = b (it
 * may alias W_out) warm-starts a later call on a nearby matrix (the next
 * HOOI step of the same mode, tensorly/decomposition/_tucker.py:194-202)
 * without re-orthonormalising it; 0 < k0 < b columns are a partial warm
 * start as for tla_sym_eig.  tla_gram_svd_block returns 0 where no block
 * exists (n <= 128: Jacobi on the whole matrix, W0 and W_out unused).
 * Workspace: tla_sym_eig_workspace_size(n). */
int tla_gram_svd_block(int n, int k);
int tla_gram_left_singular(const double* G, int n, int k, const double* W0, int k0, int dtype, void* U,
                           void* s, double* U64, double* W_out, double tol, int max_iter, int* iters_out,
                           void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Native CP-ALS sweep driver: the mode loop of
 * tensorly/decomposition/_cp.py:394-531 (parafac) run from C++.
 * The handle keeps device pointers to X, the factors (updated in place) and
 * the weights, plus a carve-up of the caller's workspace (FP64 Grams, ‖X‖²,
 * the MTTKRP buffer, dimension-tree intermediates and kernel scratch).
 *   fixed_modes: ndim flags (nullable); the last mode is never fixed.
 *   dimtree: 1 = share partial contractions between modes (orders 3, 4).
 * tla_cp_als_refresh computes ‖X‖² (with_norm) and the factor Grams; call it
 * before the first sweep and after changing factors outside the driver.
 * tla_cp_als_sweeps runs n sweeps; errors[first_slot + s] (device FP64,
 * nullable) receives the relative reconstruction error of sweep s.
 * ------------------------------------------------------------------- */
size_t tla_cp_als_workspace_size(int x_dtype, int ndim, const int64_t* shape, int rank, int precision,
                                 int dimtree);
int tla_cp_als_dimtree_supported(int x_dtype, int ndim, const int64_t* shape, int rank, int precision);
int tla_cp_als_create(int x_dtype, const void* X, int ndim, const int64_t* shape, int rank,
                      void* const* factors, void* weights, double l2_reg, int precision, int dimtree,
                      const int* fixed_modes, void* workspace, size_t workspace_bytes,
                      void** handle_out);
int tla_cp_als_dimtree_active(void* handle);
int tla_cp_als_refresh(void* handle, int with_norm, void* stream);
int tla_cp_als_set_normx2(void* handle, const double* normx2, void* stream);
int tla_cp_als_sweeps(void* handle, int n_sweeps, double* errors, int first_slot, int normalize,
                      void* stream);
/* copy state to a device buffer: 0 Grams (ndim x R x R FP64), 1 ‖X‖²,
 * 2 last <M, A> (FP64), 3 last solve status (int), 4 largest solve status
 * since the last refresh (int; 2 = a singular normal-equations matrix was
 * met, where tensorly/decomposition/_cp.py:425-427's tl.solve raises) */
int tla_cp_als_copy_state(void* handle, int which, void* dst, size_t bytes, void* stream);
void tla_cp_als_destroy(void* handle);

/* ---------------------------------------------------------------------
 * Learned extrapolation of the ALS iterates.  Replaces the fixed step
 * jump = iteration^(1/acc_pow) of the reference line search
 * (tensorly/decomposition/_cp.py:470-503) by a step chosen per iteration by
 * a 16 -> 64 -> 64 -> 1 tanh network whose input is the error history.
 *   tla_mlp_num_params: parameter count (5313), packed
 *     W1[64][16] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3[1]  (FP32).
 *   tla_mlp_forward:  out[i] = net(X[i, 0:16]) for i < n.
 *   tla_mlp_train:    `steps` Adam steps of mean squared error on minibatches
 *     of `batch` (<= 128) rows taken from `order` (device int, n entries,
 *     used cyclically); Adam moments in adam_m/adam_v, step count starts at
 *     t0 + 1; loss[step] (device, nullable) receives each minibatch loss.
 *     With grad_out non-NULL no update is made: one step over the first
 *     `batch` rows writes the exact parameter gradient there.
 *   tla_cp_als_run_extrapolated: n_sweeps sweeps of the driver with error
 *     slots errors[0..n_sweeps).  After `warmup` plain sweeps it repeats:
 *     save A_prev, one sweep, alpha = clip(2^net(features) - 1, 0, alpha_max)
 *     (alpha_fixed when mlp_params is NULL), A <- A + alpha (A - A_prev), one
 *     sweep from the extrapolated point, and keep it only if the error fell
 *     (accepted[s] = 1), else restore the pre-extrapolation factors and copy
 *     the previous error forward.  alphas[s], accepted[s] (device) describe
 *     slot s; features (nullable, n_sweeps x 16 FP32) receives the network
 *     inputs.  Workspace: tla_cp_als_extrap_workspace_size(handle) bytes.
 *   tla_extrap_features: the 16 network inputs of slot s from the error,
 *     alpha and acceptance history (the same device code the driver runs).
 * ------------------------------------------------------------------- */
int tla_mlp_num_params(void);
int tla_mlp_forward(const float* params, const float* X, int n, float* out, void* stream);
int tla_mlp_train(float* params, float* adam_m, float* adam_v, const float* X, const float* y, const int* order,
                  int n, int batch, int steps, float lr, float beta1, float beta2, float eps, int t0, float* loss,
                  float* grad_out, void* stream);
size_t tla_cp_als_extrap_workspace_size(void* handle);
int tla_cp_als_run_extrapolated(void* handle, int n_sweeps, const float* mlp_params, float alpha_fixed,
                                float alpha_max, int warmup, double* errors, int* accepted, float* alphas,
                                float* features, void* workspace, size_t workspace_bytes, void* stream);
int tla_extrap_features(const double* errors, const float* alphas, const int* accepted, int s, float* out,
                        void* stream);

/* ---------------------------------------------------------------------
 * Batched CP-ALS: n_problems independent dense 3-way float32 problems
 * (every mode size 1..64, rank 1..16) run to convergence in ONE persistent
 * launch; each workgroup pulls problems from a device queue.  Per problem
 * it is the loop of tensorly/decomposition/_cp.py:394-531 (weights fixed at
 * ones, no normalisation, tol on |Δ rec_error| as cvg_criterion
 * 'abs_rec_error', :506-521) with
 *   policy 0  plain ALS;
 *          1  the reference line search (_cp.py:333-336,401-403,448-483);
 *          2  learned (mlp_params) or fixed (alpha_fixed, mlp_params NULL)
 *             extrapolation, verified by the next sweep, after `warmup`
 *             plain sweeps; a rejected step pauses extrapolation for 1, 2,
 *             4 or 8 plain sweeps (doubling while rejections follow each
 *             other, reset by an accepted step);
 *          3  probing: as 2, but at decisions first_decision ..
 *             first_decision + max_decisions - 1 every step in
 *             probe_alphas[n_probe] is tried (one sweep each, errors to
 *             probe_err[p][d][g], features to features[p][d][16]) and the
 *             best one is taken; other decisions use the network (if
 *             mlp_params) or alpha_fixed.
 * Layout (all device): X packed, tensor p at X + x_off[p] (C-order,
 * shapes[3p..3p+2]); factors packed, problem p's A | B | C (row-major
 * I_m x ranks[p]) at factors + f_off[p], updated in place; f_prev / f_save
 * are scratch of the same size; order[] is the queue order (a permutation);
 * queue is one device int (reset by the call).  History arrays (nullable
 * for policies 0 and 1) are [n_problems][max_sweeps].  Outputs per problem:
 * sweeps, tensor passes, final relative error, extrapolations tried and
 * kept, status (0 ok, 2 singular, 3 non-finite), ticks (nullable, 100 MHz
 * wall clock).  waves: 2, 4 or 8 per workgroup; grid: workgroups; ktile:
 * the MTTKRP's reduction tile, 16, 32 or 64 reduction indices in flight
 * per wave (the same results to the bit; a launch-shape parameter for the
 * autotuner).
 * Needs no workspace beyond the arrays above and allocates nothing, so the
 * call can be captured in a HIP graph (the queue reset is a kernel).
 * tla_batched_cp_max_blocks_per_cu: resident workgroups per CU for `waves`.
 * ------------------------------------------------------------------- */
int tla_batched_cp_max_blocks_per_cu(int waves);
int tla_batched_cp_als(int n_problems, const float* X, const int64_t* x_off, const int* shapes, const int* ranks,
                       float* factors, const int64_t* f_off, float* f_prev, float* f_save, const int* order,
                       int* queue, int policy, int max_sweeps, double tol, int warmup, const float* mlp_params,
                       float alpha_fixed, float alpha_max, const float* probe_alphas, int n_probe,
                       int max_decisions, int first_decision, double* hist_err, float* hist_alpha, int* hist_acc, float* features,
                       double* probe_err, int* sweeps_out, int* passes_out, double* err_out, int* n_extrap_out,
                       int* n_accept_out, int* status_out, long long* ticks_out, int waves, int grid,
                       int ktile, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TENSORLY_AMD_H */
