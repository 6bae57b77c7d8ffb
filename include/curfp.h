/*
 * curfp.h — C ABI of the MI355X (gfx950) Rectangular Full Packed (RFP) library.
 *
 * Drop-in for the hot path of the reference header include/curfp.h:
 * same entry-point names, same enum values, same argument meaning, same
 * status codes.  The only type change is cudaStream_t -> hipStream_t; the
 * cuBLAS / cuSOLVER includes are gone because nothing here calls them.
 *
 * Conventions (identical to the reference):
 *   - every matrix is a caller-owned device pointer;
 *   - alpha, beta, value and info are HOST pointers;
 *   - every operation is enqueued on the stream bound to the handle
 *     (curfpSetStream); the default is the null stream;
 *   - spftrf is host-synchronous (it returns info through a host pointer),
 *     everything else is asynchronous on the handle's stream;
 *   - no entry point throws; all return a curfpStatus_t.
 *
 * All internal indexing is 64-bit, so n up to 2^31-1 is accepted as long as
 * the n*(n+1)/2 packed array fits in device memory (n = 65536 already has
 * n*(n+1)/2 > INT_MAX).
 */
#ifndef CURFP_H
#define CURFP_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes — reference include/curfp.h:14-20 (same values). */
typedef enum {
    CURFP_STATUS_SUCCESS          = 0,
    CURFP_STATUS_NOT_INITIALIZED  = 1,
    CURFP_STATUS_ALLOC_FAILED     = 2,
    CURFP_STATUS_INVALID_VALUE    = 3,
    CURFP_STATUS_EXECUTION_FAILED = 4
} curfpStatus_t;

/* TRANSR / TRANS — reference include/curfp.h:25-28. */
typedef enum {
    CURFP_OP_N = 0,
    CURFP_OP_T = 1
} curfpOperation_t;

/* UPLO — reference include/curfp.h:33-36. */
typedef enum {
    CURFP_FILL_MODE_LOWER = 0,
    CURFP_FILL_MODE_UPPER = 1
} curfpFillMode_t;

/* Side of curfpSsfmm / curfpDsfmm — reference include/curfp.h:38-44. */
typedef enum {
    CURFP_SIDE_LEFT  = 0,   /* C = alpha * A * B + beta * C */
    CURFP_SIDE_RIGHT = 1    /* C = alpha * B * A + beta * C */
} curfpSideMode_t;

/* Norm type of curfpSlansf / curfpDlansf — reference include/curfp.h:46-53. */
typedef enum {
    CURFP_NORM_MAX = 0,   /* max |A(i,j)|                                  */
    CURFP_NORM_ONE = 1,   /* largest absolute column sum (= infinity norm) */
    CURFP_NORM_FRO = 2    /* Frobenius norm                                */
} curfpNormType_t;

/*
 * Arithmetic used by the contractions of ssfrk and of the spftrf updates
 * (TRSM / SYRK / GEMM).  No counterpart in the reference (which always runs
 * cuBLAS FP32); the FP64 d* routines ignore it.
 *   CURFP_MATH_FP32   : FP32 inputs, v_mfma_f32_32x32x2_f32, FP32 accumulate (default).
 *   CURFP_MATH_BF16   : operands rounded to BF16, v_mfma_f32_32x32x16_bf16, FP32 accumulate.
 *   CURFP_MATH_BF16X3 : split-BF16 (x = hi + lo, hi*hi + hi*lo + lo*hi), FP32 accumulate.
 * The small diagonal-tile Cholesky and triangular inverses always run in FP32.
 */
typedef enum {
    CURFP_MATH_FP32   = 0,
    CURFP_MATH_BF16   = 1,
    CURFP_MATH_BF16X3 = 2
} curfpMathMode_t;

typedef struct curfpContext *curfpHandle_t;

/* Handle management — reference include/curfp.h:63-66, src/curfp_handle.cpp:4-51.
 * curfpSetStream differs in one respect: the handle's cached workspace is
 * shared by all of its calls, so switching to a different stream makes the
 * new stream wait for the work already queued on the old one (an event; the
 * old stream must still be valid at the switch). */
curfpStatus_t curfpCreate(curfpHandle_t *handle);
curfpStatus_t curfpDestroy(curfpHandle_t handle);
curfpStatus_t curfpSetStream(curfpHandle_t handle, hipStream_t stream);
curfpStatus_t curfpGetStream(curfpHandle_t handle, hipStream_t *stream);

/*
 * curfpSynchronize — no reference counterpart.  Waits for the handle's
 * stream and reports a failure that an earlier asynchronous call could only
 * detect on the device: a substitution sweep of spftrs (FP32) whose bounded
 * wait for an earlier block ran out, or whose block counter was shared with
 * an overlapping launch, writes NaN into its right-hand sides and marks the
 * handle.  This call, and every call that synchronises with the host anyway
 * (spftrf, spfcon, curfpSpftrsRefine), then returns
 * CURFP_STATUS_EXECUTION_FAILED once and clears the mark.
 */
curfpStatus_t curfpSynchronize(curfpHandle_t handle);

/* Polls a sweep workgroup makes for an earlier block before it gives up
 * (default 2^20, a quarter of a second or more; >= 0).  No reference
 * counterpart: a bound that keeps a corrupted progress word from hanging the
 * device. */
curfpStatus_t curfpSetSweepWaitLimit(curfpHandle_t handle, int polls);

/* Device memory currently held by the handle (cached workspace and the BF16
 * shadow of the factor); no reference counterpart. */
curfpStatus_t curfpGetWorkspaceSize(curfpHandle_t handle, size_t *bytes);

/* Math mode for the S routines (no reference counterpart). */
curfpStatus_t curfpSetMathMode(curfpHandle_t handle, curfpMathMode_t mode);
curfpStatus_t curfpGetMathMode(curfpHandle_t handle, curfpMathMode_t *mode);

/*
 * curfpSsfrk — reference include/curfp.h:79-91, src/curfp_ssfrk.cpp:135-212.
 *   trans = N: C := alpha*A*A^T + beta*C, A is n x k column-major (lda >= n)
 *   trans = T: C := alpha*A^T*A + beta*C, A is k x n column-major (lda >= k),
 *              i.e. a row-major (n, k) array.
 * C is the n x n symmetric matrix in RFP format (n*(n+1)/2 elements).
 * Quick returns as in the reference: n == 0; (alpha == 0 or k == 0) and
 * beta == 1 is a no-op; alpha == beta == 0 zero-fills C (A is never read).
 */
curfpStatus_t curfpSsfrk(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, curfpOperation_t trans,
                         int n, int k, const float *alpha, const float *A,
                         int lda, const float *beta, float *C);

/* curfpSpftrf — reference include/curfp.h:105-112, src/curfp_spftrf.cpp:182-292.
 * In-place Cholesky of the RFP matrix (uplo = L: A = L*L^T, uplo = U: A = U^T*U).
 * *info = 0 on success, > 0 = order of the first leading minor that is not
 * positive definite (the factor is then partially overwritten). */
curfpStatus_t curfpSpftrf(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, float *A, int *info);

/* curfpSpftrs — reference include/curfp.h:123-132, src/curfp_spftrs.cpp:215-302.
 * Solves A*X = B with the RFP factor from curfpSpftrf; B is n x nrhs
 * column-major (ldb >= n) and is overwritten with X. */
curfpStatus_t curfpSpftrs(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, int nrhs,
                          const float *A, float *B, int ldb);

/* curfpSpftrsRowMajor — curfpSpftrs with B stored ROW-major (n x nrhs,
 * ldb >= nrhs), the layout of the reference's Python API
 * (curfp/__init__.py:540), which otherwise transposes B into a column-major
 * copy and back (curfp/__init__.py:548, :562).  No reference C counterpart. */
curfpStatus_t curfpSpftrsRowMajor(curfpHandle_t handle, curfpOperation_t transr,
                                  curfpFillMode_t uplo, int n, int nrhs,
                                  const float *A, float *B, int ldb);

/* curfpSstrttf — reference include/curfp.h:141-149, src/curfp_strttf.cu:421-462.
 * Dense n x n ROW-MAJOR triangle (lda >= n) -> RFP. */
curfpStatus_t curfpSstrttf(curfpHandle_t handle, curfpOperation_t transr,
                           curfpFillMode_t uplo, int n, const float *A,
                           int lda, float *arf);

/* curfpSstfttr — reference include/curfp.h:197-205, src/curfp_stfttr.cu:134-175.
 * RFP -> dense n x n ROW-MAJOR triangle (lda >= n); the other triangle is
 * not written. */
curfpStatus_t curfpSstfttr(curfpHandle_t handle, curfpOperation_t transr,
                           curfpFillMode_t uplo, int n, const float *arf,
                           float *A, int lda);

/* curfpSaddToDiagonal — C entry point for the reference's pure-torch
 * add_to_diagonal (curfp/__init__.py:194-296): arf[diag] += *value for the n
 * diagonal positions of the RFP matrix, one kernel, on the handle stream.
 * The result is fl(arf + value) elementwise. */
curfpStatus_t curfpSaddToDiagonal(curfpHandle_t handle, curfpOperation_t transr,
                                  curfpFillMode_t uplo, int n,
                                  const float *value, float *arf);

/* curfpSsfmv — reference include/curfp.h:151-172, src/curfp_ssfmv.cpp:134.
 * y := alpha*A*x + beta*y with A symmetric in RFP, x and y with BLAS
 * increments incx, incy (nonzero; negative walks backwards).  alpha = 0 reads
 * neither arf nor x; beta = 0 does not read y. */
curfpStatus_t curfpSsfmv(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, int n, const float *alpha,
                         const float *arf, const float *x, int incx,
                         const float *beta, float *y, int incy);

/* curfpSpfcon — reference include/curfp.h:229-250, src/curfp_spfcon.cu:84.
 * Estimate of rcond = 1 / (||A^-1||_1 anorm) from the RFP Cholesky factor arf
 * (Hager-Higham / LAPACK ?lacn2, at most 5 iterations); anorm = ||A||_1 of
 * the matrix before factorisation.  *rcond (host) is 0 for n = 0 or
 * anorm = 0.  Synchronises the handle's stream once per iteration. */
curfpStatus_t curfpSpfcon(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, const float *arf,
                          float anorm, float *rcond);

/* curfpSsfr — reference include/curfp.h:252-270, src/curfp_ssfr.cpp:99.
 * C := alpha*x*x^T + C on the RFP matrix; x has n elements with BLAS increment
 * incx (nonzero; negative walks backwards).  alpha is a host pointer. */
curfpStatus_t curfpSsfr(curfpHandle_t handle, curfpOperation_t transr,
                        curfpFillMode_t uplo, int n, const float *alpha,
                        const float *x, int incx, float *arf);

/* curfpSsfr2 — reference include/curfp.h:272-292, src/curfp_ssfr2.cpp:94.
 * C := alpha*x*y^T + alpha*y*x^T + C. */
curfpStatus_t curfpSsfr2(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, int n, const float *alpha,
                         const float *x, int incx, const float *y, int incy,
                         float *arf);

/* curfpSsfr2k — reference include/curfp.h:294-319, src/curfp_ssfr2k.cpp:121.
 * trans = N: C := alpha*A*B^T + alpha*B*A^T + beta*C (A, B n x k, lda, ldb >= n);
 * trans = T: C := alpha*A^T*B + alpha*B^T*A + beta*C (A, B k x n, lda, ldb >= k).
 * Column-major A, B; honours curfpSetMathMode like curfpSsfrk. */
curfpStatus_t curfpSsfr2k(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, curfpOperation_t trans, int n, int k,
                          const float *alpha, const float *A, int lda,
                          const float *B, int ldb, const float *beta, float *C);

/* curfpSpftri — reference include/curfp.h:175-186, src/curfp_spftri.cu:159-388.
 * Overwrites the RFP Cholesky factor from curfpSpftrf with inv(C) in the same
 * RFP layout (L^-1 in place, then L^-T L^-1).  Asynchronous on the handle
 * stream. */
curfpStatus_t curfpSpftri(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, float *arf);

/* curfpSsfmm — reference include/curfp.h:322-348, src/curfp_ssfmm.cpp:107-306.
 * side L: C := alpha*A*B + beta*C (A m x m); side R: C := alpha*B*A + beta*C
 * (A n x n); A symmetric in RFP, B and C column-major m x n (ldb, ldc >= m),
 * B and C not overlapping.  alpha == 0: A and B are not read.  beta == 0: C is
 * not read.  One launch on FP32 MFMA (no cuBLAS SYMM + GEMM split). */
curfpStatus_t curfpSsfmm(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, curfpSideMode_t side, int m, int n,
                         const float *alpha, const float *arf, const float *B, int ldb,
                         const float *beta, float *C, int ldc);

/* curfpSlansf — reference include/curfp.h:207-227, src/curfp_slansf.cu:273-435.
 * Norm of the symmetric matrix stored in RFP (the matrix itself, not its
 * Cholesky factor); *result is a host pointer, the call synchronizes the
 * handle stream.  Unlike the reference (int-sized cuBLAS reductions,
 * INVALID_VALUE once n(n+1)/2 > INT_MAX) every size is 64-bit. */
curfpStatus_t curfpSlansf(curfpHandle_t handle, curfpNormType_t norm,
                          curfpOperation_t transr, curfpFillMode_t uplo, int n,
                          const float *arf, float *result);

/* curfpSsfrkBf16 — ssfrk with a BF16 A (raw 16-bit storage, same layout and
 * lda meaning as curfpSsfrk).  Products of BF16 values are exact in FP32, so
 * the result is the FP32-accumulated A*A^T of the BF16 matrix.  No reference
 * counterpart. */
curfpStatus_t curfpSsfrkBf16(curfpHandle_t handle, curfpOperation_t transr,
                             curfpFillMode_t uplo, curfpOperation_t trans,
                             int n, int k, const float *alpha,
                             const uint16_t *A, int lda, const float *beta,
                             float *C);

/*
 * curfpSgenResidual — FP64 residual for iterative refinement (no reference
 * counterpart):
 *   R := B - (alpha * op(A) * op(A)^T + lambda * I) * X        (all n x nrhs)
 * computed from the generator A (FP32, trans/lda as in curfpSsfrk) with FP64
 * products and FP64 accumulation.  X, B, R are FP64, column-major, with
 * leading dimensions ldx, ldb, ldr.  R may alias B.
 */
curfpStatus_t curfpSgenResidual(curfpHandle_t handle, curfpOperation_t trans,
                                int n, int k, int nrhs, const double *alpha,
                                const float *A, int lda, const double *lambda,
                                const double *X, int ldx, const double *B,
                                int ldb, double *R, int ldr);

/*
 * Kernel profiling (no reference counterpart; the reference has only
 * CUDA-event timing in its benchmark scripts).  When enabled, every kernel
 * launch of the library is bracketed by a pair of HIP events on the handle
 * stream and tagged with its algorithmic work (flops, or bytes for the
 * bandwidth kernels).  curfpGetKernelStats synchronises the stream, folds
 * the finished launches into per-class totals and returns them.
 */
typedef enum {
    CURFP_KERNEL_GEMM_F32   = 0,   /* view GEMM, FP32 MFMA                */
    CURFP_KERNEL_GEMM_BF16  = 1,   /* view GEMM, BF16 / split-BF16 MFMA   */
    CURFP_KERNEL_GEMM_F64   = 2,   /* view GEMM, FP64 MFMA                */
    CURFP_KERNEL_POTRF_TILE = 3,   /* leaf Cholesky + inverse             */
    CURFP_KERNEL_TRTRI      = 4,   /* batched leaf inverses (spftrs)      */
    CURFP_KERNEL_PACK       = 5,   /* strttf / stfttr                     */
    CURFP_KERNEL_DIAG       = 6,   /* add_to_diagonal                     */
    CURFP_KERNEL_OTHER      = 7,   /* fills, reductions, residual setup   */
    CURFP_KERNEL_TRSV       = 8,   /* spftrs substitutions                */
    CURFP_KERNEL_TRSM       = 9,   /* spftrf panel triangular solves      */
    CURFP_KERNEL_CLASSES    = 10
} curfpKernelClass_t;

curfpStatus_t curfpSetProfiling(curfpHandle_t handle, int enable);
/* Totals since the last reset for one class; reset != 0 clears them after reading. */
curfpStatus_t curfpGetKernelStats(curfpHandle_t handle, curfpKernelClass_t cls,
                                  int64_t *launches, double *flops,
                                  double *bytes, double *milliseconds, int reset);

/*
 * curfpSpftrsRefine — mixed-precision iterative refinement (no reference
 * counterpart) for
 *     C X = B,   C = alpha * op(A) * op(A)^T + lambda * I   (n x n),
 * given the FP32 RFP Cholesky factor F of C (from curfpSpftrf, in any math
 * mode) and the FP32 generator A (trans / lda as in curfpSsfrk).
 *   X_0 = F^{-1} B;  repeat  R = B - C X (FP64, from the generator),
 *   X += F^{-1} fp32(R)  (FP64 accumulation)
 * until ||R||_F <= tol * ||C||_F * ||X||_F or max_iter solves.
 * Under CURFP_MATH_BF16 (the handle's mode at this call) the solves' FP32
 * substitution sweeps take their off-diagonal products in BF16 (FP32
 * accumulation), matching the factor's BF16 accuracy; other modes solve in FP32.
 * B and X are FP64 column-major n x nrhs (ldb, ldx >= n); X need not be
 * initialised.  *iters = solves done, *resid = relative residual of the
 * returned X: the iterate of smallest residual (a step that increases the
 * residual, a sign that the factor is too inaccurate for classical
 * refinement, is undone, and the loop stops once the residual exceeds twice
 * its best).  Host-synchronous (reads the residual norm once per iteration).
 */
curfpStatus_t curfpSpftrsRefine(curfpHandle_t handle, curfpOperation_t transr,
                                curfpFillMode_t uplo, curfpOperation_t trans,
                                int n, int k, int nrhs, const double *alpha,
                                const float *A, int lda, const double *lambda,
                                const float *F, const double *B, int ldb,
                                double *X, int ldx, double tol, int max_iter,
                                int *iters, double *resid);

/*
 * curfpSpftrsNoAlloc — no counterpart in the reference.  curfpSpftrs for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc): the same FP32 sweep
 * solve, same arguments and results, with its scratch (leaf inverses and
 * look-back products) in a fixed 40 MiB per-device array that the library
 * holds from its first curfpCreate on that device until the process ends;
 * no allocation in the call.  Limits: n <= 16384 and nrhs <= 4096, else
 * CURFP_STATUS_INVALID_VALUE.  All calls on one GPU share that array, so
 * calls that overlap on one GPU (other streams or handles) race.
 */
curfpStatus_t curfpSpftrsNoAlloc(curfpHandle_t handle, curfpOperation_t transr,
                                 curfpFillMode_t uplo, int n, int nrhs,
                                 const float *A, float *B, int ldb);

/*
 * curfpSgemmStrided / curfpDgemmStrided — the library's view GEMM on dense
 * strided operands (no reference counterpart; used by the autograd backward):
 *     C(i, j) = alpha * sum_k A(i, k) * B(j, k) + beta * C(i, j)
 * with X(i, j) at X[i*x_rs + j*x_cs]; each of A, B, C needs a unit stride in
 * one direction.  beta == 0 never reads C.
 */
curfpStatus_t curfpSgemmStrided(curfpHandle_t handle, int m, int n, int k,
                                const float *alpha, const float *A, int64_t a_rs, int64_t a_cs,
                                const float *B, int64_t b_rs, int64_t b_cs,
                                const float *beta, float *C, int64_t c_rs, int64_t c_cs);
curfpStatus_t curfpDgemmStrided(curfpHandle_t handle, int m, int n, int k,
                                const double *alpha, const double *A, int64_t a_rs, int64_t a_cs,
                                const double *B, int64_t b_rs, int64_t b_cs,
                                const double *beta, double *C, int64_t c_rs, int64_t c_cs);

/* ---------------------------- FP64 twins -------------------------------- */

/* curfpDsfrk  — reference include/curfp.h:357-369, src/curfp_dssfrk.cpp:101-172. */
curfpStatus_t curfpDsfrk(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, curfpOperation_t trans,
                         int n, int k, const double *alpha, const double *A,
                         int lda, const double *beta, double *C);

/* curfpDpftrf — reference include/curfp.h:374-381, src/curfp_dpftrf.cpp:140-244. */
curfpStatus_t curfpDpftrf(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, double *A, int *info);

/* curfpDpftrs — reference include/curfp.h:386-395, src/curfp_dpftrs.cpp:117-195. */
curfpStatus_t curfpDpftrs(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, int nrhs,
                          const double *A, double *B, int ldb);

/* curfpDpftrsRowMajor — FP64 twin of curfpSpftrsRowMajor
 * (reference Python layout curfp/__init__.py:1547-1581). */
curfpStatus_t curfpDpftrsRowMajor(curfpHandle_t handle, curfpOperation_t transr,
                                  curfpFillMode_t uplo, int n, int nrhs,
                                  const double *A, double *B, int ldb);

/* curfpDstrttf — reference include/curfp.h:400-408, src/curfp_dstrttf.cu:111-152. */
curfpStatus_t curfpDstrttf(curfpHandle_t handle, curfpOperation_t transr,
                           curfpFillMode_t uplo, int n, const double *A,
                           int lda, double *arf);

/* curfpDstfttr — reference include/curfp.h:441-449, src/curfp_dstfttr.cu:111-152. */
curfpStatus_t curfpDstfttr(curfpHandle_t handle, curfpOperation_t transr,
                           curfpFillMode_t uplo, int n, const double *arf,
                           double *A, int lda);

/* curfpDaddToDiagonal — FP64 twin of curfpSaddToDiagonal
 * (reference curfp/__init__.py:272-296 accepts float64 too). */
curfpStatus_t curfpDaddToDiagonal(curfpHandle_t handle, curfpOperation_t transr,
                                  curfpFillMode_t uplo, int n,
                                  const double *value, double *arf);

/* curfpDsfmv — reference include/curfp.h:410-424.  FP64 twin of curfpSsfmv. */
curfpStatus_t curfpDsfmv(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, int n, const double *alpha,
                         const double *arf, const double *x, int incx,
                         const double *beta, double *y, int incy);

/* curfpDpfcon — reference include/curfp.h:464-476.  FP64 twin of curfpSpfcon. */
curfpStatus_t curfpDpfcon(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, const double *arf,
                          double anorm, double *rcond);

/* curfpDsfr / curfpDsfr2 / curfpDsfr2k — reference include/curfp.h:478-525.
 * FP64 twins of curfpSsfr / curfpSsfr2 / curfpSsfr2k. */
curfpStatus_t curfpDsfr(curfpHandle_t handle, curfpOperation_t transr,
                        curfpFillMode_t uplo, int n, const double *alpha,
                        const double *x, int incx, double *arf);
curfpStatus_t curfpDsfr2(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, int n, const double *alpha,
                         const double *x, int incx, const double *y, int incy,
                         double *arf);
curfpStatus_t curfpDsfr2k(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, curfpOperation_t trans, int n, int k,
                          const double *alpha, const double *A, int lda,
                          const double *B, int ldb, const double *beta, double *C);

/* curfpDpftri — reference include/curfp.h:427-436.  FP64 twin of curfpSpftri. */
curfpStatus_t curfpDpftri(curfpHandle_t handle, curfpOperation_t transr,
                          curfpFillMode_t uplo, int n, double *arf);

/* curfpDsfmm — reference include/curfp.h:528-544, src/curfp_dsfmm.cpp:78-219.  FP64 twin of
 * curfpSsfmm. */
curfpStatus_t curfpDsfmm(curfpHandle_t handle, curfpOperation_t transr,
                         curfpFillMode_t uplo, curfpSideMode_t side, int m, int n,
                         const double *alpha, const double *arf, const double *B, int ldb,
                         const double *beta, double *C, int ldc);

/* curfpDlansf — reference include/curfp.h:451-462, src/curfp_dlansf.cu:189.
 * FP64 twin of curfpSlansf (64-bit sizes, host result, synchronizing). */
curfpStatus_t curfpDlansf(curfpHandle_t handle, curfpNormType_t norm,
                          curfpOperation_t transr, curfpFillMode_t uplo, int n,
                          const double *arf, double *result);

#ifdef __cplusplus
}
#endif

#endif /* CURFP_H */
