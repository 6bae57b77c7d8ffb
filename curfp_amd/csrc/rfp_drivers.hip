/*
 * rfp_drivers.hip — host drivers behind the C ABI (include/curfp.h).
 *
 * Every routine works on the three strided views of the RFP rectangle
 * (rfp_geom in rfp_internal.h) and issues hand-written gfx950 kernels on the
 * handle's stream:
 *
 *   ssfrk  : three view GEMMs (lower-triangular SYRK on v11 and v22, full
 *            GEMM on v21) — the reference's 2x cublasSsyrk + cublasSgemm
 *            (src/curfp_ssfrk.cpp:193-209);
 *   spftrf : LAPACK's RFP sequence POTRF(v11) -> TRSM(v21) -> SYRK(v22) ->
 *            POTRF(v22) (src/curfp_spftrf.cpp:256-279), with the two POTRFs
 *            and the TRSM themselves blocked recursively down to leaf tiles
 *            whose Cholesky + inverse run in one workgroup; every update
 *            above the leaves is an MFMA GEMM;
 *   spftrs : forward  L11 y1 = b1, b2 -= L21 y1, L22 y2 = b2,
 *            backward L22^T x2 = y2, b1 -= L21^T x2, L11^T x1 = b1
 *            (src/curfp_spftrs.cpp:258-299), each triangular solve blocked
 *            recursively with the leaf solves done as GEMMs with the leaf
 *            inverses.
 *
 * spftrf accumulates `info` on the device and reads it back once, at the end
 * (the reference synchronises twice, src/curfp_spftrf.cpp:258-260, :280-282).
 * Workspace (leaf inverses, split-K partials) is cached in the handle instead
 * of being allocated per call (src/curfp_spftrf.cpp:231-232).
 */
#include "rfp_internal.h"
#include "trsm_panel_shape.h"
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

/* ------------------------------------------------------------------------ */
/* Handle                                                                    */
/* ------------------------------------------------------------------------ */

static curfpStatus_t st_of(hipError_t e)
{
    if (e == hipSuccess) return CURFP_STATUS_SUCCESS;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CURFP_STATUS_ALLOC_FAILED;
    if (e == hipErrorInvalidValue) return CURFP_STATUS_INVALID_VALUE;
    return CURFP_STATUS_EXECUTION_FAILED;
}

#define RFP_TRY(expr)                                   \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return _e;                \
    } while (0)

#define RFP_CHECK_HANDLE(h)                             \
    do {                                                \
        if (!(h)) return CURFP_STATUS_NOT_INITIALIZED;  \
    } while (0)

static void noalloc_scratch_init(int device);

extern "C" curfpStatus_t curfpCreate(curfpHandle_t *handle)
{
    if (!handle) return CURFP_STATUS_INVALID_VALUE;
    curfpContext *ctx = new (std::nothrow) curfpContext();
    if (!ctx) return CURFP_STATUS_ALLOC_FAILED;
    ctx->stream = nullptr;
    ctx->math = CURFP_MATH_FP32;
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
    ctx->d_info = nullptr;
    ctx->sweep_sync = nullptr;
    ctx->sweep_epoch = 0;
    ctx->sweep_wait_limit = SWEEP_WAIT_LIMIT_DEFAULT;
    ctx->fence = nullptr;
    ctx->prof = false;
    ctx->sh = nullptr;
    ctx->sh_bytes = 0;
    ctx->blas = nullptr;
    ctx->blas_off = blas_native_only();
    for (RfpKernelStat &st : ctx->prof_stats) st = RfpKernelStat{0, 0.0, 0.0, 0.0};
    if (hipGetDevice(&ctx->device) != hipSuccess) { delete ctx; return CURFP_STATUS_NOT_INITIALIZED; }
    noalloc_scratch_init(ctx->device);
    if (hipMalloc((void **)&ctx->d_info, sizeof(int)) != hipSuccess) { delete ctx; return CURFP_STATUS_ALLOC_FAILED; }
    if (hipMalloc(&ctx->sweep_sync, SWEEP_SYNC_BYTES) != hipSuccess ||
        hipMemset(ctx->sweep_sync, 0, SWEEP_SYNC_BYTES) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) {
        (void)hipFree(ctx->d_info);
        if (ctx->sweep_sync) (void)hipFree(ctx->sweep_sync);
        delete ctx;
        return CURFP_STATUS_ALLOC_FAILED;
    }
    *handle = ctx;
    return CURFP_STATUS_SUCCESS;
}

extern "C" curfpStatus_t curfpDestroy(curfpHandle_t handle)
{
    RFP_CHECK_HANDLE(handle);
    (void)hipStreamSynchronize(handle->stream);
    for (const RfpProfRecord &r : handle->prof_pending) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    for (hipEvent_t e : handle->prof_pool) (void)hipEventDestroy(e);
    if (handle->fence) (void)hipEventDestroy(handle->fence);
    blas_destroy(handle->blas);
    if (handle->ws) (void)hipFree(handle->ws);
    if (handle->sh) (void)hipFree(handle->sh);
    if (handle->d_info) (void)hipFree(handle->d_info);
    if (handle->sweep_sync) (void)hipFree(handle->sweep_sync);
    delete handle;
    return CURFP_STATUS_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Kernel profiling                                                          */
/* ------------------------------------------------------------------------ */

static hipEvent_t prof_event(curfpHandle_t h)
{
    if (!h->prof_pool.empty()) {
        hipEvent_t e = h->prof_pool.back();
        h->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

/* Runs `launch` (which enqueues one kernel on h->stream); when profiling is
 * on, brackets it with events and records its algorithmic work. */
template <typename F>
static hipError_t timed(curfpHandle_t h, int cls, double flops, double bytes, F &&launch)
{
    if (!h->prof) return launch();
    hipEvent_t e0 = prof_event(h), e1 = prof_event(h);
    if (!e0 || !e1) return launch();
    (void)hipEventRecord(e0, h->stream);
    hipError_t e = launch();
    (void)hipEventRecord(e1, h->stream);
    if (e == hipSuccess) {
        h->prof_pending.push_back(RfpProfRecord{cls, flops, bytes, e0, e1});
    } else {
        /* nothing of this class ran (a failed launch, or a library that
         * declined the problem and handed it back) */
        h->prof_pool.push_back(e0);
        h->prof_pool.push_back(e1);
    }
    return e;
}

static hipError_t prof_flush(curfpHandle_t h)
{
    if (h->prof_pending.empty()) return hipSuccess;
    RFP_TRY(hipStreamSynchronize(h->stream));
    for (const RfpProfRecord &r : h->prof_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
            RfpKernelStat &st = h->prof_stats[r.cls];
            st.launches += 1;
            st.flops += r.flops;
            st.bytes += r.bytes;
            st.ms += ms;
        }
        h->prof_pool.push_back(r.e0);
        h->prof_pool.push_back(r.e1);
    }
    h->prof_pending.clear();
    return hipSuccess;
}

extern "C" curfpStatus_t curfpSetProfiling(curfpHandle_t handle, int enable)
{
    RFP_CHECK_HANDLE(handle);
    hipError_t e = prof_flush(handle);
    if (e != hipSuccess) return st_of(e);
    handle->prof = enable != 0;
    for (RfpKernelStat &st : handle->prof_stats) st = RfpKernelStat{0, 0.0, 0.0, 0.0};
    return CURFP_STATUS_SUCCESS;
}

extern "C" curfpStatus_t curfpGetKernelStats(curfpHandle_t handle, curfpKernelClass_t cls, int64_t *launches,
                                             double *flops, double *bytes, double *milliseconds, int reset)
{
    RFP_CHECK_HANDLE(handle);
    if ((int)cls < 0 || (int)cls >= CURFP_KERNEL_CLASSES) return CURFP_STATUS_INVALID_VALUE;
    hipError_t e = prof_flush(handle);
    if (e != hipSuccess) return st_of(e);
    const RfpKernelStat &st = handle->prof_stats[cls];
    if (launches) *launches = st.launches;
    if (flops) *flops = st.flops;
    if (bytes) *bytes = st.bytes;
    if (milliseconds) *milliseconds = st.ms;
    if (reset) handle->prof_stats[cls] = RfpKernelStat{0, 0.0, 0.0, 0.0};
    return CURFP_STATUS_SUCCESS;
}

/* The handle's workspace, BF16 shadow, info word and sweep counters are
 * reused by every call, so a stream switch must not let the new stream's
 * work overtake what the old stream still has queued: the new stream waits
 * on an event recorded on the old one (the stream being left must still be
 * valid).  The reference allocates its workspace per call on the stream
 * instead (src/curfp_spftrf.cpp:231-232, 288-291) and only stores the
 * stream (src/curfp_handle.cpp:37-43). */
extern "C" curfpStatus_t curfpSetStream(curfpHandle_t handle, hipStream_t stream)
{
    RFP_CHECK_HANDLE(handle);
    if (stream != handle->stream) {
        if (!handle->fence && hipEventCreateWithFlags(&handle->fence, hipEventDisableTiming) != hipSuccess) {
            handle->fence = nullptr;
            return CURFP_STATUS_EXECUTION_FAILED;
        }
        if (hipEventRecord(handle->fence, handle->stream) != hipSuccess ||
            hipStreamWaitEvent(stream, handle->fence, 0) != hipSuccess)
            return CURFP_STATUS_EXECUTION_FAILED;
    }
    handle->stream = stream;
    return CURFP_STATUS_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Device-side failures of asynchronous calls                                */
/* ------------------------------------------------------------------------ */

/* A substitution sweep that gave up waiting (or found its ticket counter
 * shared with another launch) writes NaN and sets the error word in the
 * handle's sync area.  Entry points that synchronise with the host anyway
 * (spftrf's info, the refinement's norms, spfcon's records,
 * curfpSynchronize) enqueue a copy of that word with their own read-back
 * and, when it is set, clear the sync area and return EXECUTION_FAILED —
 * the status the reference gives a runtime failure
 * (src/curfp_internal.h:80-85). */
static hipError_t sweep_err_fetch(curfpHandle_t h, int *host)
{
    *host = 0;
    return hipMemcpyAsync(host, sweep_err_word(h->sweep_sync), sizeof(int), hipMemcpyDeviceToHost, h->stream);
}

static hipError_t sweep_err_check(curfpHandle_t h, int word)
{
    if (!word) return hipSuccess;
    /* progress words, counters and the error word back to zero for the next sweep */
    (void)launch_zero_f64((double *)h->sweep_sync, SWEEP_SYNC_BYTES / (int)sizeof(double), h->stream);
    return hipErrorLaunchFailure;
}

extern "C" curfpStatus_t curfpSynchronize(curfpHandle_t handle)
{
    RFP_CHECK_HANDLE(handle);
    int word = 0;
    hipError_t e = sweep_err_fetch(handle, &word);
    if (e == hipSuccess) e = hipStreamSynchronize(handle->stream);
    if (e == hipSuccess) e = sweep_err_check(handle, word);
    return st_of(e);
}

extern "C" curfpStatus_t curfpSetSweepWaitLimit(curfpHandle_t handle, int polls)
{
    RFP_CHECK_HANDLE(handle);
    if (polls < 0) return CURFP_STATUS_INVALID_VALUE;
    handle->sweep_wait_limit = polls;
    return CURFP_STATUS_SUCCESS;
}

extern "C" curfpStatus_t curfpGetStream(curfpHandle_t handle, hipStream_t *stream)
{
    RFP_CHECK_HANDLE(handle);
    if (!stream) return CURFP_STATUS_INVALID_VALUE;
    *stream = handle->stream;
    return CURFP_STATUS_SUCCESS;
}

extern "C" curfpStatus_t curfpGetWorkspaceSize(curfpHandle_t handle, size_t *bytes)
{
    RFP_CHECK_HANDLE(handle);
    if (!bytes) return CURFP_STATUS_INVALID_VALUE;
    *bytes = handle->ws_bytes + handle->sh_bytes;
    return CURFP_STATUS_SUCCESS;
}

extern "C" curfpStatus_t curfpSetMathMode(curfpHandle_t handle, curfpMathMode_t mode)
{
    RFP_CHECK_HANDLE(handle);
    if (mode != CURFP_MATH_FP32 && mode != CURFP_MATH_BF16 && mode != CURFP_MATH_BF16X3)
        return CURFP_STATUS_INVALID_VALUE;
    handle->math = mode;
    return CURFP_STATUS_SUCCESS;
}

extern "C" curfpStatus_t curfpGetMathMode(curfpHandle_t handle, curfpMathMode_t *mode)
{
    RFP_CHECK_HANDLE(handle);
    if (!mode) return CURFP_STATUS_INVALID_VALUE;
    *mode = handle->math;
    return CURFP_STATUS_SUCCESS;
}

hipError_t rfp_workspace(curfpHandle_t h, size_t bytes, void **out)
{
    if (bytes == 0) bytes = 256;
    if (h->ws_bytes < bytes) {
        /* the old buffer may still be in use by queued work */
        RFP_TRY(hipStreamSynchronize(h->stream));
        if (h->ws) { (void)hipFree(h->ws); h->ws = nullptr; h->ws_bytes = 0; }
        const size_t sz = (bytes + (1u << 20) - 1) & ~((size_t)(1u << 20) - 1);
        RFP_TRY(hipMalloc(&h->ws, sz));
        h->ws_bytes = sz;
    }
    *out = h->ws;
    return hipSuccess;
}

/* ------------------------------------------------------------------------ */
/* Precision dispatch                                                        */
/* ------------------------------------------------------------------------ */

template <typename T> struct Prec;
template <> struct Prec<float> {
    static hipError_t gemm(const GemmArgs &g, RfpMath m, hipStream_t s) { return launch_gemm_f32(g, m, s); }
    static hipError_t potrf_tile(float *a, RfpView v, int m, int *info, int base, float *inv, hipStream_t s)
    { return launch_potrf_tile_f32(a, v, m, info, base, inv, s); }
    static hipError_t trtri(const float *a, RfpView v, int64_t m, float *inv, hipStream_t s)
    { return launch_trtri_tiles_f32(a, v, m, inv, s); }
    static hipError_t trtri2(const float *a, RfpView v1, int64_t m1, float *inv1, RfpView v2, int64_t m2, float *inv2,
                             hipStream_t s)
    { return launch_trtri_tiles2_f32(a, v1, m1, inv1, v2, m2, inv2, s); }
    static hipError_t trttf(const RfpGeom &g, bool lo, const float *A, int64_t lda, float *arf, hipStream_t s)
    { return launch_trttf_f32(g, lo, A, lda, arf, s); }
    static hipError_t tfttr(const RfpGeom &g, bool lo, const float *arf, float *A, int64_t lda, hipStream_t s)
    { return launch_tfttr_f32(g, lo, arf, A, lda, s); }
    static hipError_t add_diag(const RfpGeom &g, float v, float *arf, hipStream_t s)
    { return launch_add_diag_f32(g, v, arf, s); }
    static constexpr bool f64 = false;
};
template <> struct Prec<double> {
    static hipError_t gemm(const GemmArgs &g, RfpMath, hipStream_t s) { return launch_gemm_f64(g, s); }
    static hipError_t potrf_tile(double *a, RfpView v, int m, int *info, int base, double *inv, hipStream_t s)
    { return launch_potrf_tile_f64(a, v, m, info, base, inv, s); }
    static hipError_t trtri(const double *a, RfpView v, int64_t m, double *inv, hipStream_t s)
    { return launch_trtri_tiles_f64(a, v, m, inv, s); }
    static hipError_t trtri2(const double *a, RfpView v1, int64_t m1, double *inv1, RfpView v2, int64_t m2,
                             double *inv2, hipStream_t s)
    { return launch_trtri_tiles2_f64(a, v1, m1, inv1, v2, m2, inv2, s); }
    static hipError_t trttf(const RfpGeom &g, bool lo, const double *A, int64_t lda, double *arf, hipStream_t s)
    { return launch_trttf_f64(g, lo, A, lda, arf, s); }
    static hipError_t tfttr(const RfpGeom &g, bool lo, const double *arf, double *A, int64_t lda, hipStream_t s)
    { return launch_tfttr_f64(g, lo, arf, A, lda, s); }
    static hipError_t add_diag(const RfpGeom &g, double v, double *arf, hipStream_t s)
    { return launch_add_diag_f64(g, v, arf, s); }
    static constexpr bool f64 = true;
};

/* Lower-triangular outputs of at least this order are cut into diagonal
 * blocks of order below it (for the hand-written SYRK tiles, one batched
 * launch) and the row panels left of them (for the vendor library).
 * Measured on the n = 65536 chains (FP64 n = 32768): BF16 operands 8192
 * (4096: +5.6 ms per BF16+IR step, 16384: +1.6 ms), FP32 4096 (8192:
 * +3.7 ms, 16384: +12 ms), FP64 8192 (4096: +5.7 ms, 16384: +3.2 ms). */
static int64_t lower_split_min(const GemmArgs &g, bool f64)
{
    if (g.bf16_rows || g.a_bf16) return 8192;
    return f64 ? 8192 : 4096;
}

/* Algorithmic flops of one view GEMM: 2*M*N*K, or M*(M+1)*K for a
 * lower-triangular (SYRK) output. */
static double gemm_flops(const GemmArgs &g)
{
    return g.lower ? (double)g.M * (double)(g.M + 1) * (double)g.K : 2.0 * g.M * (double)g.N * (double)g.K;
}

/* Algorithmic HBM bytes of one view GEMM: each operand read once (A alone
 * when B is the same view, as in a SYRK), C written once and read too when
 * beta != 0, plus the BF16 shadow written by the epilogue.  Split-K partials
 * are not algorithmic and are not counted. */
template <typename T>
static double gemm_bytes(const GemmArgs &g, RfpMath m)
{
    const bool bf16_in = g.a_bf16 || g.bf16_rows;
    const bool x3 = g.bf16_rows && m == RFP_MATH_BF16X3;
    const double ea = bf16_in ? 2.0 * (x3 ? 2 : 1) : (g.a_f32 ? 4.0 : (double)sizeof(T));
    const double eb = bf16_in ? 2.0 * (x3 ? 2 : 1) : (g.b_f32 ? 4.0 : (double)sizeof(T));
    const bool same = g.a == g.b && g.va.off == g.vb.off && g.va.rs == g.vb.rs && g.va.cs == g.vb.cs;
    const double out = g.lower ? (double)g.M * (double)(g.M + 1) / 2.0 : (double)g.M * (double)g.N;
    double bytes = (double)g.M * g.K * ea + (same ? 0.0 : (double)g.N * g.K * eb);
    bytes += out * (double)sizeof(T) * (g.beta != 0.0 ? 2.0 : 1.0);
    if (g.c_sh_hi) bytes += out * 2.0 * (g.c_sh_lo ? 2.0 : 1.0);
    return bytes;
}

template <typename T>
static hipError_t run_gemm(curfpHandle_t h, const GemmArgs &g, RfpMath m)
{
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    const int cls = Prec<T>::f64 ? CURFP_KERNEL_GEMM_F64
                                 : ((m == RFP_MATH_F32 && !g.a_bf16) ? CURFP_KERNEL_GEMM_F32 : CURFP_KERNEL_GEMM_BF16);
    if (g.lower && g.batch <= 1 && g.splits <= 1 && !g.inplace && !g.c_sh_hi && g.M >= lower_split_min(g, Prec<T>::f64) &&
        (Prec<T>::f64 ? !g.a_f32 : (g.bf16_rows || (m == RFP_MATH_F32 && !g.a_bf16)))) {
        /* A large lower-triangular output: nb diagonal blocks of order S (the
         * last one m_last) and the row panels left of them.  The panels are
         * plain GEMMs for the vendor library; the diagonal triangles are one
         * batched launch of the hand-written SYRK tiles (separate launches of
         * a few hundred tiles each would leave most of the chip idle in their
         * last round). */
        const int64_t smin = lower_split_min(g, Prec<T>::f64);
        int64_t nb = 2;
        while ((g.M + nb - 1) / nb >= smin) nb *= 2;
        const int64_t S = ((g.M + nb - 1) / nb + 255) / 256 * 256;
        nb = (g.M + S - 1) / S;
        auto panel = [&](int64_t z) {
            GemmArgs r = g;
            r.lower = false;
            r.M = std::min<int64_t>(S, g.M - z * S);
            r.N = z * S;
            r.va = g.va.sub(z * S, 0);
            r.vc = g.vc.sub(z * S, 0);
            return r;
        };
        bool lib = nb > 1;
        for (int64_t z = 1; z < nb && lib; ++z) lib = blas_route(h, panel(z), m, Prec<T>::f64) != 0;
        if (lib) {
            for (int64_t z = 1; z < nb; ++z) RFP_TRY(run_gemm<T>(h, panel(z), m));
            GemmArgs t = g;
            t.M = t.N = S;
            t.batch = (int)nb;
            t.a_bs = S * g.va.rs;
            t.b_bs = S * g.vb.rs;
            t.c_bs = S * (g.vc.rs + g.vc.cs);
            t.m_last = g.M - (nb - 1) * S;
            GemmArgs last = g;
            last.M = last.N = t.m_last;
            const double fl = (double)(nb - 1) * gemm_flops(t) + gemm_flops(last);
            const double by = (double)(nb - 1) * gemm_bytes<T>(t, m) + gemm_bytes<T>(last, m);
            return timed(h, cls, fl, by, [&] { return Prec<T>::gemm(t, m, h->stream); });
        }
    }
    if (const int route = blas_route(h, g, m, Prec<T>::f64)) {
        /* hipErrorNotSupported: the library has no plan for this problem (no
         * algorithm, or its state could not be created) — the hand-written
         * kernels take it */
        const hipError_t e =
            timed(h, cls, gemm_flops(g), gemm_bytes<T>(g, m), [&] { return blas_gemm(h, g, m, route); });
        if (e != hipErrorNotSupported) return e;
    }
    return timed(h, cls, gemm_flops(g), gemm_bytes<T>(g, m), [&] { return Prec<T>::gemm(g, m, h->stream); });
}

static RfpMath math_of(curfpHandle_t h)
{
    switch (h->math) {
    case CURFP_MATH_BF16: return RFP_MATH_BF16;
    case CURFP_MATH_BF16X3: return RFP_MATH_BF16X3;
    default: return RFP_MATH_F32;
    }
}

/* ------------------------------------------------------------------------ */
/* ssfrk                                                                     */
/* ------------------------------------------------------------------------ */

template <typename T>
static curfpStatus_t sfrk_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo,
                               curfpOperation_t trans, int n, int k, const T *alpha,
                               const void *A, int lda, const T *beta, T *C, bool a_bf16)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || k < 0 || lda < 1) return CURFP_STATUS_INVALID_VALUE;
    if (!alpha || !beta) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) || (trans != CURFP_OP_N && trans != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    const T a = *alpha, b = *beta;
    if ((a == T(0) || k == 0) && b == T(1)) return CURFP_STATUS_SUCCESS;
    if (!C) return CURFP_STATUS_INVALID_VALUE;
    const int64_t nt = (int64_t)n * (n + 1) / 2;
    if (a == T(0) && b == T(0)) return st_of(launch_fill_zero(C, (size_t)nt * sizeof(T), h->stream));
    const int64_t K = (a == T(0)) ? 0 : k;     /* alpha == 0: A is never read */
    if (K > 0 && !A) return CURFP_STATUS_INVALID_VALUE;

    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    /* row i of op(A) (n x k): trans = N -> A[i + p*lda], trans = T -> A[p + i*lda] */
    const RfpView av = (trans == CURFP_OP_N) ? RfpView{0, 1, lda} : RfpView{0, lda, 1};
    const RfpView a1 = av, a2 = av.sub(g.n1, 0);
    /* a BF16 input is exact in BF16: its split-BF16 low part is zero */
    const RfpMath m = a_bf16 ? RFP_MATH_BF16 : math_of(h);
    /* BF16 rows with k contiguous and 16-byte aligned go to the row-block kernel */
    const bool rows = a_bf16 && trans == CURFP_OP_T && lda % 8 == 0 && ((uintptr_t)A % 16) == 0;

    struct Blk { RfpView c, x, y; int64_t M, N; bool lower; } blk[3] = {
        {g.v11, a1, a1, g.n1, g.n1, true},
        {g.v21, a2, a1, g.n2, g.n1, false},
        {g.v22, a2, a2, g.n2, g.n2, true},
    };
    /* FP32 A in a BF16 mode on a large update: op(A) converted once to BF16
     * row blocks (k padded to a multiple of 8 with zeros) for the row-block
     * kernels, instead of converting every operand tile as it is staged */
    if (!a_bf16 && !Prec<T>::f64 && m != RFP_MATH_F32 && n >= 2048 && K >= 32) {
        const int64_t Kp = (K + 7) / 8 * 8;
        const bool x3 = m == RFP_MATH_BF16X3;
        const size_t bytes = ((size_t)n * Kp * sizeof(unsigned short) + 255) & ~(size_t)255;
        void *ws = nullptr;
        hipError_t e = rfp_workspace(h, bytes * (x3 ? 2 : 1), &ws);
        if (e != hipSuccess) return st_of(e);
        unsigned short *P = (unsigned short *)ws;
        unsigned short *Pl = x3 ? (unsigned short *)((char *)ws + bytes) : nullptr;
        e = timed(h, CURFP_KERNEL_PACK, 0.0, (double)n * K * 4.0 + (double)n * Kp * (x3 ? 4.0 : 2.0), [&] {
            return launch_to_bf16_rows((const float *)A, av, n, K, Kp, Kp, P, Pl, h->stream);
        });
        if (e != hipSuccess) return st_of(e);
        const int64_t r0[3] = {0, g.n1, g.n1}, c0[3] = {0, 0, g.n1};
        for (int q = 0; q < 3; ++q) {
            const Blk &bk = blk[q];
            if (bk.M <= 0 || bk.N <= 0) continue;
            GemmArgs ga = gemm_args();
            ga.a = P; ga.a_lo = Pl; ga.va = RfpView{r0[q] * Kp, Kp, 1};
            ga.b = P; ga.b_lo = Pl; ga.vb = RfpView{c0[q] * Kp, Kp, 1};
            ga.bf16_rows = true;
            ga.c = C; ga.vc = bk.c;
            ga.M = bk.M; ga.N = bk.N; ga.K = Kp;
            ga.alpha = (double)a; ga.beta = (double)b;
            ga.lower = bk.lower;
            e = run_gemm<T>(h, ga, m);
            if (e != hipSuccess) return st_of(e);
        }
        return CURFP_STATUS_SUCCESS;
    }
    for (const Blk &bk : blk) {
        if (bk.M <= 0 || bk.N <= 0) continue;
        GemmArgs ga = gemm_args();
        ga.a = A; ga.va = bk.x;
        ga.b = A; ga.vb = bk.y;
        ga.c = C; ga.vc = bk.c;
        ga.M = bk.M; ga.N = bk.N; ga.K = K;
        ga.alpha = (double)a; ga.beta = (double)b;
        ga.lower = bk.lower;
        ga.a_bf16 = a_bf16;
        ga.bf16_rows = rows;
        hipError_t e = run_gemm<T>(h, ga, m);
        if (e != hipSuccess) return st_of(e);
    }
    return CURFP_STATUS_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* ssfr / ssfr2 / ssfr2k (and D)                                             */
/* ------------------------------------------------------------------------ */

/*
 * C := alpha (X Y^T + Y X^T) + beta C on the RFP views (Y == nullptr: C :=
 * alpha X X^T + beta C), with X, Y n x k given by their row views (row i of
 * X at xv.sub(i, 0)).  Reference src/curfp_ssfr2k.cpp:121, src/curfp_ssfr2.cpp:94,
 * src/curfp_ssfr.cpp:99: per block of the RFP matrix (v11 lower, v21 full,
 * v22 lower) one or two GEMMs, the second accumulating onto the first.
 * alpha == 0 does not read X or Y; beta == 0 does not read C.
 */
template <typename T>
static curfpStatus_t sfr2k_core(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int64_t n,
                                int64_t k, T a, const T *X, RfpView xv, const T *Y, RfpView yv, T b, T *C)
{
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if ((a == T(0) || k == 0) && b == T(1)) return CURFP_STATUS_SUCCESS;
    if (!C) return CURFP_STATUS_INVALID_VALUE;
    const int64_t nt = n * (n + 1) / 2;
    if ((a == T(0) || k == 0) && b == T(0)) return st_of(launch_fill_zero(C, (size_t)nt * sizeof(T), h->stream));
    const int64_t K = (a == T(0)) ? 0 : k;
    const bool two = Y != nullptr;
    if (K > 0 && (!X || (two && !Y))) return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const RfpMath m = Prec<T>::f64 ? RFP_MATH_F64 : math_of(h);
    /* BF16 modes on a large update: op(A), op(B) converted once to BF16 row
     * blocks P = [X | Y] and Q = [Y | X] (each half k padded to a multiple of
     * 8 with zeros), so P Q^T = X Y^T + Y X^T is one GEMM of depth 2 Kp per
     * view on the BF16 row-block kernels (direct global -> LDS loads): C is
     * read and written once, and no operand tile is converted while staged */
    const bool convert = !Prec<T>::f64 && m != RFP_MATH_F32 && K > 0 && n >= 2048 && K >= 32;
    if (convert) {
        const int64_t Kp = (K + 7) / 8 * 8, ld = 2 * Kp;
        const bool x3 = m == RFP_MATH_BF16X3;
        const size_t blk = ((size_t)n * ld * sizeof(unsigned short) + 255) & ~(size_t)255;
        void *ws = nullptr;
        hipError_t e = rfp_workspace(h, blk * (x3 ? 4 : 2), &ws);
        if (e != hipSuccess) return st_of(e);
        unsigned short *P = (unsigned short *)ws, *Q = (unsigned short *)((char *)ws + blk);
        unsigned short *Pl = x3 ? (unsigned short *)((char *)ws + 2 * blk) : nullptr;
        unsigned short *Ql = x3 ? (unsigned short *)((char *)ws + 3 * blk) : nullptr;
        const T *Yo = two ? Y : X;
        const RfpView yo = two ? yv : xv;
        struct Cv { const T *src; RfpView v; unsigned short *hi, *lo; } cv[4] = {
            {X, xv, P, Pl}, {Yo, yo, P + Kp, Pl ? Pl + Kp : nullptr},
            {Yo, yo, Q, Ql}, {X, xv, Q + Kp, Ql ? Ql + Kp : nullptr},
        };
        for (const Cv &c : cv) {
            e = timed(h, CURFP_KERNEL_PACK, 0.0, (double)n * K * 4.0 + (double)n * Kp * (x3 ? 4.0 : 2.0), [&] {
                return launch_to_bf16_rows((const float *)c.src, c.v, n, K, Kp, ld, c.hi, c.lo, h->stream);
            });
            if (e != hipSuccess) return st_of(e);
        }
        struct Blk { RfpView c; int64_t r0, c0, M, N; bool lower; } blk3[3] = {
            {g.v11, 0, 0, g.n1, g.n1, true},
            {g.v21, g.n1, 0, g.n2, g.n1, false},
            {g.v22, g.n1, g.n1, g.n2, g.n2, true},
        };
        for (const Blk &bk : blk3) {
            if (bk.M <= 0 || bk.N <= 0) continue;
            GemmArgs ga = gemm_args();
            ga.c = C; ga.vc = bk.c;
            ga.M = bk.M; ga.N = bk.N; ga.K = ld;
            ga.lower = bk.lower;
            ga.bf16_rows = true;
            ga.a = P; ga.a_lo = Pl; ga.va = RfpView{bk.r0 * ld, ld, 1};
            ga.b = Q; ga.b_lo = Ql; ga.vb = RfpView{bk.c0 * ld, ld, 1};
            /* a rank-k update (no Y) is P = [X | X], Q = [X | X] at half alpha */
            ga.alpha = two ? (double)a : 0.5 * (double)a;
            ga.beta = (double)b;
            e = run_gemm<T>(h, ga, m);
            if (e != hipSuccess) return st_of(e);
        }
        return CURFP_STATUS_SUCCESS;
    }
    const RfpView x1 = xv, x2 = xv.sub(g.n1, 0), y1 = yv, y2 = yv.sub(g.n1, 0);
    struct Blk { RfpView c, xr, xc, yr, yc; int64_t M, N; bool lower; } blk[3] = {
        {g.v11, x1, x1, y1, y1, g.n1, g.n1, true},
        {g.v21, x2, x1, y2, y1, g.n2, g.n1, false},
        {g.v22, x2, x2, y2, y2, g.n2, g.n2, true},
    };
    for (const Blk &bk : blk) {
        if (bk.M <= 0 || bk.N <= 0) continue;
        GemmArgs ga = gemm_args();
        ga.c = C; ga.vc = bk.c;
        ga.M = bk.M; ga.N = bk.N; ga.K = K;
        ga.lower = bk.lower;
        /* X_rows Y_cols^T (or X_rows X_cols^T), then Y_rows X_cols^T */
        ga.a = X; ga.va = bk.xr;
        ga.b = two ? Y : X; ga.vb = two ? bk.yc : bk.xc;
        ga.alpha = (double)a; ga.beta = (double)b;
        hipError_t e = run_gemm<T>(h, ga, m);
        if (e == hipSuccess && two && K > 0) {
            ga.a = Y; ga.va = bk.yr;
            ga.b = X; ga.vb = bk.xc;
            ga.beta = 1.0;
            e = run_gemm<T>(h, ga, m);
        }
        if (e != hipSuccess) return st_of(e);
    }
    return CURFP_STATUS_SUCCESS;
}

/* row view of an n-vector with BLAS increment inc (inc < 0 walks backwards
 * from element (n - 1) |inc|) as an n x 1 matrix */
static inline RfpView vec_view(int64_t n, int inc)
{
    return inc > 0 ? RfpView{0, inc, 1} : RfpView{(n - 1) * (int64_t)(-inc), inc, 1};
}

template <typename T>
static curfpStatus_t sfr_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const T *alpha,
                              const T *x, int incx, const T *y, int incy, bool two, T *arf)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || !alpha || incx == 0 || (two && incy == 0)) return CURFP_STATUS_INVALID_VALUE;
    return sfr2k_core<T>(h, transr, uplo, n, 1, *alpha, x, vec_view(n, incx), two ? y : nullptr,
                         two ? vec_view(n, incy) : RfpView{0, 1, 1}, T(1), arf);
}

template <typename T>
static curfpStatus_t sfr2k_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo,
                                curfpOperation_t trans, int n, int k, const T *alpha, const T *A, int lda,
                                const T *B, int ldb, const T *beta, T *C)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || k < 0 || lda < 1 || ldb < 1 || !alpha || !beta) return CURFP_STATUS_INVALID_VALUE;
    if (trans != CURFP_OP_N && trans != CURFP_OP_T) return CURFP_STATUS_INVALID_VALUE;
    if (trans == CURFP_OP_N ? (n > 0 && (lda < n || ldb < n)) : (k > 0 && (lda < k || ldb < k)))
        return CURFP_STATUS_INVALID_VALUE;
    /* row i of op(A) (n x k): trans = N -> A[i + p*lda], trans = T -> A[p + i*lda] */
    const RfpView av = trans == CURFP_OP_N ? RfpView{0, 1, lda} : RfpView{0, lda, 1};
    const RfpView bv = trans == CURFP_OP_N ? RfpView{0, 1, ldb} : RfpView{0, ldb, 1};
    return sfr2k_core<T>(h, transr, uplo, n, k, *alpha, A, av, B, bv, *beta, C);
}

/* ------------------------------------------------------------------------ */
/* spftrf                                                                    */
/* ------------------------------------------------------------------------ */

/*
 * In BF16 / split-BF16 mode the factorisation keeps a BF16 shadow of the
 * factor: three dense row-major blocks (L11, L21, L22; leading dimensions
 * rounded up to 8 elements so rows are 16-byte aligned) holding bf16(L) (and
 * bf16(L - bf16(L)) for split-BF16).  A block of L is final when a leaf
 * triangular solve writes it; that GEMM's epilogue writes the shadow too.
 * Every BF16 update (the TRSM updates and the SYRKs) reads only final
 * blocks, so it reads its operands from the shadow: 2 bytes per element, k
 * contiguous, no conversion in the hot loop.
 */
template <typename T>
struct Factor {
    curfpHandle_t h;
    T *a;          /* RFP array */
    T *inv;        /* leaf inverse slots */
    int64_t leaf;
    RfpMath math;  /* arithmetic of the SYRK / GEMM updates */
    unsigned short *sh_hi, *sh_lo;  /* BF16 shadow (null in FP32 / FP64 mode) */

    hipError_t gemm(RfpView c, const T *x, RfpView vx, const T *y, RfpView vy, int64_t M, int64_t N,
                    int64_t K, double alpha, double beta, bool lower, int inplace, RfpMath m,
                    const RfpView *shadow = nullptr)
    {
        GemmArgs g = gemm_args();
        g.a = x; g.va = vx;
        g.b = y; g.vb = vy;
        g.c = a; g.vc = c;
        g.M = M; g.N = N; g.K = K;
        g.alpha = alpha; g.beta = beta;
        g.lower = lower;
        g.inplace = inplace;
        if (shadow && sh_hi) {
            g.c_sh_hi = sh_hi;
            g.c_sh_lo = sh_lo;
            g.vcs = *shadow;
        }
        return run_gemm<T>(h, g, m);
    }

    /* C -= X Y^T with X, Y read from the BF16 shadow */
    hipError_t gemm_shadow(RfpView c, RfpView sx, RfpView sy, int64_t M, int64_t N, int64_t K, bool lower)
    {
        GemmArgs g = gemm_args();
        g.a = sh_hi; g.va = sx; g.a_lo = sh_lo;
        g.b = sh_hi; g.vb = sy; g.b_lo = sh_lo;
        g.bf16_rows = true;
        g.c = a; g.vc = c;
        g.M = M; g.N = N; g.K = K;
        g.alpha = -1.0; g.beta = 1.0;
        g.lower = lower;
        return run_gemm<T>(h, g, math);
    }

    /* B (mb x m) := B * L^{-T}, L = lower-triangular view (m x m) whose
     * leaf inverses start at `slot`; Bs / Ls are the shadow views */
    hipError_t trsm_rt(RfpView B, RfpView Bs, int64_t mb, RfpView L, RfpView Ls, int64_t m, int64_t slot)
    {
        if (m <= 0 || mb <= 0) return hipSuccess;
        if (!Prec<T>::f64 && m <= trsm_panel_max()) {
            /* FP32: the last levels of the recursion in one launch, each
             * workgroup substituting a slab of rows through all leaves. */
            const double fl = (double)mb * (double)m * (double)m;
            const double by = (double)mb * m * 4.0 * 2.0 + (double)m * m * 4.0 + (sh_hi ? (double)mb * m * 2.0 : 0.0);
            return timed(h, CURFP_KERNEL_TRSM, fl, by, [&] {
                return launch_trsm_panel_f32((float *)a, B, mb, L, m, (const float *)(inv + slot * leaf * leaf),
                                             sh_hi, sh_lo, Bs, math == RFP_MATH_BF16, h->stream);
            });
        }
        if (Prec<T>::f64 && m <= trsm_panel64_max()) {
            /* FP64: the same one-launch panel over 64-row leaves */
            const double fl = (double)mb * (double)m * (double)m;
            const double by = (double)mb * m * 8.0 * 2.0 + (double)m * m * 8.0;
            return timed(h, CURFP_KERNEL_TRSM, fl, by, [&] {
                return launch_trsm_panel_f64((double *)a, B, mb, L, m, (const double *)(inv + slot * leaf * leaf),
                                             h->stream);
            });
        }
        if (m <= leaf)
            return gemm(B, a, B, inv + slot * leaf * leaf, RfpView{0, leaf, 1}, mb, m, m, 1.0, 0.0,
                        false, 1, RFP_MATH_F32, &Bs);
        const int64_t m1 = leaf_split(m, leaf);
        RFP_TRY(trsm_rt(B, Bs, mb, L, Ls, m1, slot));
        if (sh_hi)
            RFP_TRY(gemm_shadow(B.sub(0, m1), Bs, Ls.sub(m1, 0), mb, m - m1, m1, false));
        else
            RFP_TRY(gemm(B.sub(0, m1), a, B, a, L.sub(m1, 0), mb, m - m1, m1, -1.0, 1.0, false, 0, math));
        return trsm_rt(B.sub(0, m1), Bs.sub(0, m1), mb, L.sub(m1, m1), Ls.sub(m1, m1), m - m1, slot + m1 / leaf);
    }

    hipError_t potrf2(RfpView V11, RfpView S11, RfpView V21, RfpView S21, RfpView V22, RfpView S22, int64_t m1,
                      int64_t m2, int64_t goff, int64_t slot1, int64_t slot2)
    {
        RFP_TRY(potrf(V11, S11, m1, goff, slot1));
        RFP_TRY(trsm_rt(V21, S21, m2, V11, S11, m1, slot1));
        if (m1 > 0 && m2 > 0) {
            if (sh_hi) RFP_TRY(gemm_shadow(V22, S21, S21, m2, m2, m1, true));
            else RFP_TRY(gemm(V22, a, V21, a, V21, m2, m2, m1, -1.0, 1.0, true, 0, math));
        }
        return potrf(V22, S22, m2, goff + m1, slot2);
    }

    hipError_t potrf(RfpView V, RfpView S, int64_t m, int64_t goff, int64_t slot)
    {
        if (m <= 0) return hipSuccess;
        if (m <= leaf)
            return timed(h, CURFP_KERNEL_POTRF_TILE, (double)m * m * m / 3.0, 0.0, [&] {
                return Prec<T>::potrf_tile(a, V, (int)m, h->d_info, (int)goff, inv + slot * leaf * leaf, h->stream);
            });
        if (!Prec<T>::f64 && m <= potrf_pair_max() && math != RFP_MATH_BF16X3)
            /* FP32 / BF16: the two leaves, the panel and the update between
             * them in one workgroup (four launches otherwise) */
            return timed(h, CURFP_KERNEL_POTRF_TILE, (double)m * m * m / 3.0, 0.0, [&] {
                return launch_potrf_pair_f32((float *)a, V, (int)m, h->d_info, (int)goff,
                                             (float *)(inv + slot * leaf * leaf), sh_hi, sh_lo, S,
                                             math == RFP_MATH_BF16, h->stream);
            });
        const int64_t m1 = leaf_split(m, leaf);
        return potrf2(V, S, V.sub(m1, 0), S.sub(m1, 0), V.sub(m1, m1), S.sub(m1, m1), m1, m - m1, goff, slot,
                      slot + m1 / leaf);
    }
};

hipError_t rfp_shadow(curfpHandle_t h, size_t bytes, void **out)
{
    if (bytes == 0) bytes = 256;
    if (h->sh_bytes < bytes) {
        RFP_TRY(hipStreamSynchronize(h->stream));
        if (h->sh) { (void)hipFree(h->sh); h->sh = nullptr; h->sh_bytes = 0; }
        const size_t sz = (bytes + (1u << 20) - 1) & ~((size_t)(1u << 20) - 1);
        RFP_TRY(hipMalloc(&h->sh, sz));
        h->sh_bytes = sz;
    }
    *out = h->sh;
    return hipSuccess;
}

template <typename T>
static curfpStatus_t pftrf_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                T *A, int *info)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || !info) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    *info = 0;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if (!A) return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const int64_t leaf = rfp_leaf(Prec<T>::f64);
    const int64_t s1 = (g.n1 + leaf - 1) / leaf, s2 = (g.n2 + leaf - 1) / leaf;
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, (size_t)(s1 + s2) * leaf * leaf * sizeof(T), &ws);
    if (e != hipSuccess) return st_of(e);
    e = launch_set_int(h->d_info, INT_MAX, h->stream);
    if (e != hipSuccess) return st_of(e);
    Factor<T> f{h, A, (T *)ws, leaf, Prec<T>::f64 ? RFP_MATH_F64 : math_of(h), nullptr, nullptr};
    const ShadowGeom sg = shadow_geom(g);
    if (!Prec<T>::f64 && f.math != RFP_MATH_F32) {
        const bool x3 = f.math == RFP_MATH_BF16X3;
        void *sh = nullptr;
        e = rfp_shadow(h, (size_t)sg.elems * 2 * (x3 ? 2 : 1), &sh);
        if (e != hipSuccess) return st_of(e);
        f.sh_hi = (unsigned short *)sh;
        f.sh_lo = x3 ? f.sh_hi + sg.elems : nullptr;
    }
    e = f.potrf2(g.v11, sg.s11, g.v21, sg.s21, g.v22, sg.s22, g.n1, g.n2, 0, 0, s1);
    if (e != hipSuccess) return st_of(e);
    int hinfo = INT_MAX, herr = 0;
    e = hipMemcpyAsync(&hinfo, h->d_info, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = sweep_err_fetch(h, &herr);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = sweep_err_check(h, herr);
    if (e != hipSuccess) return st_of(e);
    *info = (hinfo == INT_MAX) ? 0 : hinfo;
    return CURFP_STATUS_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* spftrs                                                                    */
/* ------------------------------------------------------------------------ */

template <typename T>
struct Solver {
    curfpHandle_t h;
    const T *a;
    T *b;
    T *inv;
    int64_t leaf, nrhs;

    T *partial;          /* split-K scratch, solve_partial_elems() elements */

    hipError_t gemm(RfpView c, const T *x, RfpView vx, const T *y, RfpView vy, int64_t M, int64_t K,
                    double alpha, double beta, int inplace)
    {
        GemmArgs g = gemm_args();
        g.a = x; g.va = vx;
        g.b = y; g.vb = vy;
        g.c = b; g.vc = c;
        g.M = M; g.N = nrhs; g.K = K;
        g.alpha = alpha; g.beta = beta;
        g.inplace = inplace;
        if (!inplace && partial) {
            /* the updates are tall-and-skinny (N = nrhs) with long K: split K
             * until there are ~4 workgroups per CU */
            const int64_t tiles = ((M + 127) / 128) * ((nrhs + 63) / 64);
            int64_t sp = (1024 + tiles - 1) / tiles;
            sp = std::min<int64_t>(sp, std::min<int64_t>(32, K / 512));
            while (sp > 1 && sp * M * nrhs > solve_partial_elems()) --sp;
            if (sp > 1) {
                g.splits = (int)sp;
                g.partial = partial;
            }
        }
        return run_gemm<T>(h, g, RFP_MATH_F32);
    }

    static int64_t solve_partial_elems() { return (int64_t)8 << 20; }

    /* Size of the diagonal blocks solved in one kernel: 512-row block solves
     * over the leaf inverses (128-row FP32 leaves, 64-row FP64 leaves). */
    int64_t block() const { return Prec<T>::f64 ? (int64_t)block_solve64_rows() : (int64_t)block_solve_rows(); }

    hipError_t leaf_solve(RfpView L, int64_t m, RfpView B, int64_t slot, bool trans)
    {
        return timed(h, CURFP_KERNEL_TRSV, 2.0 * (double)m * m * nrhs / 2.0, 0.0, [&] {
            if (Prec<T>::f64)
                return launch_block_solve_f64((const double *)a, L, (int)m, (const double *)(inv + slot * leaf * leaf),
                                              (double *)b, B, (int)nrhs, trans, h->stream);
            return launch_block_solve_f32((const float *)a, L, (int)m, (const float *)(inv + slot * leaf * leaf),
                                          (float *)b, B, (int)nrhs, trans, h->stream);
        });
    }

    /* L Y = B in place, L lower (m x m), B (m x nrhs) */
    hipError_t lower(RfpView L, int64_t m, RfpView B, int64_t slot)
    {
        if (m <= 0) return hipSuccess;
        if (m <= block()) return leaf_solve(L, m, B, slot, false);
        const int64_t m1 = leaf_split(m, block());
        RFP_TRY(lower(L, m1, B, slot));
        RFP_TRY(gemm(B.sub(m1, 0), a, L.sub(m1, 0), b, B.t(), m - m1, m1, -1.0, 1.0, 0));
        return lower(L.sub(m1, m1), m - m1, B.sub(m1, 0), slot + m1 / leaf);
    }

    /* L^T X = B in place */
    hipError_t lower_t(RfpView L, int64_t m, RfpView B, int64_t slot)
    {
        if (m <= 0) return hipSuccess;
        if (m <= block()) return leaf_solve(L, m, B, slot, true);
        const int64_t m1 = leaf_split(m, block());
        RFP_TRY(lower_t(L.sub(m1, m1), m - m1, B.sub(m1, 0), slot + m1 / leaf));
        RFP_TRY(gemm(B, a, L.sub(m1, 0).t(), b, B.sub(m1, 0).t(), m1, m - m1, -1.0, 1.0, 0));
        return lower_t(L, m1, B, slot);
    }
};

static size_t pftrs_inv_elems(const RfpGeom &g, bool f64)
{
    const int64_t leaf = rfp_leaf(f64);
    const int64_t s1 = (g.n1 + leaf - 1) / leaf, s2 = (g.n2 + leaf - 1) / leaf;
    return (size_t)(s1 + s2) * leaf * leaf;
}

/* leaf inverses, split-K partials, and (FP32) the sweep's four products per
 * leaf */
static size_t pftrs_scratch_bytes(const RfpGeom &g, bool f64)
{
    const size_t inv = pftrs_inv_elems(g, f64);
    return ((f64 ? inv : 5 * inv) + (size_t)(8 << 20)) * (f64 ? sizeof(double) : sizeof(float));
}

/* Forward + backward substitution on B (view bv, n x nrhs) with the RFP
 * factor A; `inv` holds pftrs_scratch_bytes of scratch for the leaf inverses.
 * `prepared`: the inverses (and the sweep's products) in `inv` are already
 * those of A from an earlier call with the same nrhs path.  `bf16`: BF16
 * off-diagonal products in the FP32 sweeps (column-major B only). */
template <typename T>
static hipError_t pftrs_core(curfpHandle_t h, const RfpGeom &g, const T *A, T *B, RfpView bv, int64_t nrhs, T *inv,
                             bool prepared = false, bool bf16 = false, float *pm_at = nullptr)
{
    const int64_t leaf = rfp_leaf(Prec<T>::f64);
    const int64_t s1 = (g.n1 + leaf - 1) / leaf;
    if (!prepared) {
        /* every diagonal leaf's inverse, both views in one launch */
        RFP_TRY(timed(h, CURFP_KERNEL_TRTRI, (double)g.n * leaf * leaf / 3.0, 0.0, [&] {
            return Prec<T>::trtri2(A, g.v11, g.n1, inv, g.v22, g.n2, inv + s1 * leaf * leaf, h->stream);
        }));
    }
    if (!Prec<T>::f64 && nrhs <= 64 * SWEEP_MAX_GROUPS) {
        /* FP32: one sweep launch per direction over all 128-row blocks; the
         * look-back products follow the leaf inverses and the split-K
         * partials unless the caller places them (pm_at) */
        float *pm = pm_at ? pm_at : (float *)(inv + pftrs_inv_elems(g, false)) + ((size_t)8 << 20);
        const double nb = (double)((g.n1 + leaf - 1) / leaf + (g.n2 + leaf - 1) / leaf);
        if (!prepared) {
            RFP_TRY(timed(h, CURFP_KERNEL_TRTRI, 4.0 * nb * 2.0 * leaf * leaf * leaf, 0.0, [&] {
                return launch_sweep_pmat_f32((const float *)A, g, (const float *)inv, pm, h->stream);
            }));
        }
        const double fl = (double)g.n * (double)g.n * (double)nrhs;
        const bool bf = bf16 && bv.rs == 1;
        for (int dir = 0; dir < 2; ++dir) {
            const unsigned long long ep = ++h->sweep_epoch;
            RFP_TRY(timed(h, CURFP_KERNEL_TRSV, fl, 0.0, [&] {
                return launch_sweep_solve_f32((const float *)A, g, (const float *)inv, pm, (float *)B, bv,
                                              (int)nrhs, dir == 1, bf, h->sweep_sync, ep, h->sweep_wait_limit,
                                              h->stream);
            }));
        }
        return hipSuccess;
    }
    Solver<T> sv{h, A, B, inv, leaf, nrhs, inv + pftrs_inv_elems(g, Prec<T>::f64)};
    const RfpView b2 = bv.sub(g.n1, 0);
    /* forward: L y = b */
    RFP_TRY(sv.lower(g.v11, g.n1, bv, 0));
    if (g.n1 > 0 && g.n2 > 0) RFP_TRY(sv.gemm(b2, A, g.v21, B, bv.t(), g.n2, g.n1, -1.0, 1.0, 0));
    RFP_TRY(sv.lower(g.v22, g.n2, b2, s1));
    /* backward: L^T x = y */
    RFP_TRY(sv.lower_t(g.v22, g.n2, b2, s1));
    if (g.n1 > 0 && g.n2 > 0) RFP_TRY(sv.gemm(bv, A, g.v21.t(), B, b2.t(), g.n1, g.n2, -1.0, 1.0, 0));
    return sv.lower_t(g.v11, g.n1, bv, 0);
}

template <typename T>
static curfpStatus_t pftrs_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                int nrhs, const T *A, T *B, int ldb, bool row_major)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || nrhs < 0 || ldb < 1) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0 || nrhs == 0) return CURFP_STATUS_SUCCESS;
    if (!A || !B || ldb < (row_major ? nrhs : n)) return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, pftrs_scratch_bytes(g, Prec<T>::f64), &ws);
    if (e != hipSuccess) return st_of(e);
    /* B(i, c): column-major n x nrhs (the reference ABI) or row-major */
    const RfpView bv = row_major ? RfpView{0, ldb, 1} : RfpView{0, 1, ldb};
    return st_of(pftrs_core<T>(h, g, A, B, bv, nrhs, (T *)ws));
}

/* ------------------------------------------------------------------------ */
/* spftrs without device allocation                                          */
/* ------------------------------------------------------------------------ */

/*
 * For callers that cannot allocate (HIP graph capture, serving loops): the
 * FP32 sweep solve with its scratch — the leaf inverses and the sweep's four
 * look-back products per leaf — in one fixed device array instead of the
 * handle's workspace.  At most NOALLOC_MAX_N rows: (n1/128 + n2/128) <= 128
 * leaves, 5 * 128 * 128 * 128 floats = 40 MiB, held on every device the
 * library is initialised on until the process ends.  Every call on a device
 * shares that device's array, so calls that overlap on one GPU (two streams,
 * two handles) race.  The sweep's progress words are the handle's own
 * (allocated by curfpCreate); the producers of the inverses and products
 * write their whole tiles, so nothing in the array needs clearing.
 */
constexpr int64_t NOALLOC_MAX_N = 16384;
constexpr int64_t NOALLOC_MAX_LEAVES = NOALLOC_MAX_N / 128;
constexpr size_t NOALLOC_SCRATCH_FLOATS = (size_t)5 * NOALLOC_MAX_LEAVES * 128 * 128;
static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];

/* device address of `scratch` on each device, looked up once per device by
 * curfpCreate with that device current (the host-side name is not a device
 * address) */
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static bool g_scratch_done[NOALLOC_MAX_DEVICES];

static void noalloc_scratch_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = e == hipSuccess ? (float *)ptr : nullptr;
    g_scratch_done[device] = true;
}

static float *noalloc_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

extern "C" curfpStatus_t curfpSpftrsNoAlloc(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                            int nrhs, const float *A, float *B, int ldb)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || nrhs < 0 || ldb < 1) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0 || nrhs == 0) return CURFP_STATUS_SUCCESS;
    if (!A || !B || ldb < n) return CURFP_STATUS_INVALID_VALUE;
    /* the sizes the fixed scratch holds, and the sweep path's width */
    if (n > NOALLOC_MAX_N || nrhs > 64 * SWEEP_MAX_GROUPS) return CURFP_STATUS_INVALID_VALUE;
    float *scr = noalloc_scratch(h->device);
    if (!scr) return CURFP_STATUS_EXECUTION_FAILED;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    float *pm = scr + pftrs_inv_elems(g, false);
    return st_of(pftrs_core<float>(h, g, A, B, RfpView{0, 1, ldb}, nrhs, scr, false, false, pm));
}

/* ------------------------------------------------------------------------ */
/* spfcon: reciprocal condition estimate from the RFP Cholesky factor        */
/* ------------------------------------------------------------------------ */

static hipError_t pfcon_fill(float *x, int64_t n, int mode, int64_t j, hipStream_t s)
{ return launch_pfcon_fill_f32(x, n, mode, j, s); }
static hipError_t pfcon_fill(double *x, int64_t n, int mode, int64_t j, hipStream_t s)
{ return launch_pfcon_fill_f64(x, n, mode, j, s); }
static hipError_t pfcon_sign(float *x, float *g, int64_t n, hipStream_t s) { return launch_pfcon_sign_f32(x, g, n, s); }
static hipError_t pfcon_sign(double *x, double *g, int64_t n, hipStream_t s) { return launch_pfcon_sign_f64(x, g, n, s); }
static hipError_t pfcon_stats(const float *x, const float *g, int64_t n, int64_t j, PfconStats *o, hipStream_t s)
{ return launch_pfcon_stats_f32(x, g, n, j, o, s); }
static hipError_t pfcon_stats(const double *x, const double *g, int64_t n, int64_t j, PfconStats *o, hipStream_t s)
{ return launch_pfcon_stats_f64(x, g, n, j, o, s); }

/*
 * Hager-Higham 1-norm estimate of inv(A) (LAPACK ?lacn2 with A^-T = A^-1, at
 * most 5 iterations, the decision sequence of reference
 * src/curfp_spfcon.cu:139-250), then rcond = (1 / est) / anorm.  Every
 * product inv(A) x is one forward + backward substitution on the device
 * iterate; the leaf inverses (and, in FP32, the sweep's look-back products)
 * are built by the first solve and reused by the rest.
 */
template <typename T>
static curfpStatus_t pfcon_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                const T *arf, T anorm, T *rcond)
{
    RFP_CHECK_HANDLE(h);
    if (!rcond) return CURFP_STATUS_INVALID_VALUE;
    *rcond = T(0);
    if (n == 0 || anorm == T(0)) return CURFP_STATUS_SUCCESS;
    if (n < 0 || !arf) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const size_t inv_bytes = (pftrs_scratch_bytes(g, Prec<T>::f64) + 255) & ~(size_t)255;
    const size_t vec_bytes = ((size_t)n * sizeof(T) + 255) & ~(size_t)255;
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, inv_bytes + 2 * vec_bytes + 256, &ws);
    if (e != hipSuccess) return st_of(e);
    T *inv = (T *)ws;
    T *x = (T *)((char *)ws + inv_bytes);
    T *isgn = (T *)((char *)x + vec_bytes);
    PfconStats *dst = (PfconStats *)((char *)isgn + vec_bytes);
    hipStream_t s = h->stream;
    bool prepared = false;
    auto solve = [&]() -> hipError_t {
        RFP_TRY(pftrs_core<T>(h, g, arf, x, RfpView{0, 1, n}, 1, inv, prepared));
        prepared = true;
        return hipSuccess;
    };
    PfconStats st;
    auto stats = [&](bool dot, int64_t j) -> hipError_t {
        int herr = 0;
        RFP_TRY(pfcon_stats(x, dot ? isgn : nullptr, n, j, dst, s));
        RFP_TRY(hipMemcpyAsync(&st, dst, sizeof(st), hipMemcpyDeviceToHost, s));
        RFP_TRY(sweep_err_fetch(h, &herr));
        RFP_TRY(hipStreamSynchronize(s));
        return sweep_err_check(h, herr);
    };
    constexpr int ITMAX = 5;
    auto run = [&](T &est) -> hipError_t {
        /* x = inv(A) (1/n, ..., 1/n) */
        RFP_TRY(pfcon_fill(x, n, 0, 0, s));
        RFP_TRY(solve());
        RFP_TRY(stats(false, -1));
        est = (T)st.asum;
        if (n == 1) return hipSuccess;
        /* x = inv(A) sign(x) */
        RFP_TRY(pfcon_sign(x, isgn, n, s));
        RFP_TRY(solve());
        RFP_TRY(stats(false, -1));
        int64_t j = st.imax;
        int iter = 2;
        for (;;) {
            /* x = inv(A) e_j */
            RFP_TRY(pfcon_fill(x, n, 1, j, s));
            RFP_TRY(solve());
            RFP_TRY(stats(true, -1));
            est = (T)st.asum;
            if (st.dot == st.asum) break;               /* sign pattern repeated */
            RFP_TRY(pfcon_sign(x, isgn, n, s));
            RFP_TRY(solve());
            RFP_TRY(stats(false, j));
            if (st.xj == st.xmax || iter >= ITMAX) break;
            j = st.imax;
            ++iter;
        }
        /* alternating-sign vector: est = max(est, 2 |inv(A) x|_1 / (3 n)) */
        RFP_TRY(pfcon_fill(x, n, 2, 0, s));
        RFP_TRY(solve());
        RFP_TRY(stats(false, -1));
        const T temp = (T)(2.0 * st.asum / (3.0 * (double)n));
        if (temp > est) est = temp;
        return hipSuccess;
    };
    T est = T(0);
    e = run(est);
    if (e != hipSuccess) return st_of(e);
    if (est > T(0)) *rcond = (T(1) / est) / anorm;
    return CURFP_STATUS_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* pack / unpack / diagonal                                                  */
/* ------------------------------------------------------------------------ */

template <typename T>
static curfpStatus_t trttf_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                const T *A, int lda, T *arf)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || lda < n) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if (!A || !arf) return CURFP_STATUS_INVALID_VALUE;
    const bool lower = uplo == CURFP_FILL_MODE_LOWER;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, lower);
    const double bytes = 2.0 * (double)n * (n + 1) / 2.0 * sizeof(T);
    return st_of(timed(h, CURFP_KERNEL_PACK, 0.0, bytes, [&] { return Prec<T>::trttf(g, lower, A, lda, arf, h->stream); }));
}

template <typename T>
static curfpStatus_t tfttr_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                const T *arf, T *A, int lda)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || lda < n) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if (!A || !arf) return CURFP_STATUS_INVALID_VALUE;
    const bool lower = uplo == CURFP_FILL_MODE_LOWER;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, lower);
    const double bytes = 2.0 * (double)n * (n + 1) / 2.0 * sizeof(T);
    return st_of(timed(h, CURFP_KERNEL_PACK, 0.0, bytes, [&] { return Prec<T>::tfttr(g, lower, arf, A, lda, h->stream); }));
}

template <typename T>
static curfpStatus_t add_diag_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                   const T *value, T *arf)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || !value) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if (!arf) return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const double bytes = 2.0 * (double)n * sizeof(T);
    return st_of(timed(h, CURFP_KERNEL_DIAG, 0.0, bytes, [&] { return Prec<T>::add_diag(g, *value, arf, h->stream); }));
}

/* ------------------------------------------------------------------------ */
/* spftri / dpftri                                                           */
/* ------------------------------------------------------------------------ */

/*
 * inv(C) = L^-T L^-1 from the RFP factor, in place (reference
 * src/curfp_spftri.cu:159-388; LAPACK's ?pftri: the triangular inverse, then
 * the product).  On the lower-factor views L = [L11 0; L21 L22]:
 *   X11 = inv(L11);  L21 := -L21 X11;  X22 = inv(L22);  L21 := X22 L21;
 *   C11 = X11^T X11 + X21^T X21;  C21 = X22^T X21;  C22 = X22^T X22.
 * TRTRI, TRMM and LAUUM are halved recursively down to leaves; the
 * off-diagonal pieces are view GEMMs.  Every leaf operand of the recursion is
 * an inverted diagonal tile not yet overwritten by LAUUM, so the inverses of
 * all diagonal tiles are computed up front by the batched leaf-inverse kernel
 * (dense, row-major, zero above the diagonal) and serve directly as the
 * leaf TRMM / LAUUM operands; `slot` indexes them (v11's leaves, then v22's).
 */
template <typename T>
struct Inverter {
    curfpHandle_t h;
    T *a;          /* RFP array */
    T *inv;        /* leaf x leaf inverse of diagonal tile `slot` at inv + slot*leaf*leaf */
    int64_t leaf;

    const T *tile(int64_t slot) const { return inv + slot * leaf * leaf; }
    /* C (view c of the RFP array) = alpha X Y^T + beta C */
    hipError_t gemm(RfpView c, const T *x, RfpView vx, const T *y, RfpView vy, int64_t M, int64_t N, int64_t K,
                    double alpha, double beta, bool lower, int inplace)
    {
        if (K <= 0 && beta == 1.0) return hipSuccess;
        GemmArgs g = gemm_args();
        g.a = x; g.va = vx;
        g.b = y; g.vb = vy;
        g.c = a; g.vc = c;
        g.M = M; g.N = N; g.K = K;
        g.alpha = alpha; g.beta = beta;
        g.lower = lower;
        g.inplace = inplace;
        return run_gemm<T>(h, g, Prec<T>::f64 ? RFP_MATH_F64 : RFP_MATH_F32);
    }
    /* B (mb x m) := alpha B X, X = inverted lower m x m (diagonal leaves at slot ..) */
    hipError_t trmm_right(RfpView B, int64_t mb, RfpView t, int64_t m, double alpha, int64_t slot)
    {
        if (m <= 0 || mb <= 0) return hipSuccess;
        if (m <= leaf)   /* C = alpha B Y^T, Y(j, k) = X(k, j); C aliases B */
            return gemm(B, a, B, tile(slot), RfpView{0, 1, leaf}, mb, m, m, alpha, 0.0, false, 1);
        const int64_t m1 = leaf_split(m, leaf), m2 = m - m1;
        RFP_TRY(trmm_right(B, mb, t, m1, alpha, slot));                                            /* B1 := B1 X11 */
        RFP_TRY(gemm(B, a, B.sub(0, m1), a, t.sub(m1, 0).t(), mb, m1, m2, alpha, 1.0, false, 0));  /* B1 += B2 X21 */
        return trmm_right(B.sub(0, m1), mb, t.sub(m1, m1), m2, alpha, slot + m1 / leaf);           /* B2 := B2 X22 */
    }
    /* B (m x nb) := alpha op(X) B, X inverted lower m x m, op(X) = X or X^T */
    hipError_t trmm_left(RfpView B, int64_t nb, RfpView t, int64_t m, bool trans, double alpha, int64_t slot)
    {
        if (m <= 0 || nb <= 0) return hipSuccess;
        if (m <= leaf) { /* C = alpha op(X) Y^T with Y = B^T; C aliases B */
            const RfpView wx = trans ? RfpView{0, 1, leaf} : RfpView{0, leaf, 1};
            return gemm(B, tile(slot), wx, a, B.t(), m, nb, m, alpha, 0.0, false, 2);
        }
        const int64_t m1 = leaf_split(m, leaf), m2 = m - m1, s2 = slot + m1 / leaf;
        const RfpView B2 = B.sub(m1, 0);
        if (!trans) {
            RFP_TRY(trmm_left(B2, nb, t.sub(m1, m1), m2, false, alpha, s2));                       /* B2 := X22 B2 */
            RFP_TRY(gemm(B2, a, t.sub(m1, 0), a, B.t(), m2, nb, m1, alpha, 1.0, false, 0));        /* B2 += X21 B1 */
            return trmm_left(B, nb, t, m1, false, alpha, slot);                                    /* B1 := X11 B1 */
        }
        RFP_TRY(trmm_left(B, nb, t, m1, true, alpha, slot));                                        /* B1 := X11^T B1 */
        RFP_TRY(gemm(B, a, t.sub(m1, 0).t(), a, B2.t(), m1, nb, m2, alpha, 1.0, false, 0));        /* B1 += X21^T B2 */
        return trmm_left(B2, nb, t.sub(m1, m1), m2, true, alpha, s2);                               /* B2 := X22^T B2 */
    }
    /* off-diagonal blocks of T := inv(T), T lower m x m (the diagonal leaves
     * are already inverted in `inv`) */
    hipError_t trtri(RfpView t, int64_t m, int64_t slot)
    {
        if (m <= leaf) return hipSuccess;
        const int64_t m1 = leaf_split(m, leaf), m2 = m - m1, s2 = slot + m1 / leaf;
        RFP_TRY(trtri(t, m1, slot));
        RFP_TRY(trmm_right(t.sub(m1, 0), m2, t, m1, -1.0, slot));                                  /* T21 := -T21 X11 */
        RFP_TRY(trtri(t.sub(m1, m1), m2, s2));
        return trmm_left(t.sub(m1, 0), m1, t.sub(m1, m1), m2, false, 1.0, s2);                      /* T21 := X22 T21 */
    }
    /* T := lower(X^T X) in place, X the inverted lower m x m */
    hipError_t lauum(RfpView t, int64_t m, int64_t slot)
    {
        if (m <= 0) return hipSuccess;
        if (m <= leaf)   /* C = Y Y^T (lower) with Y = X^T, Y(i, k) = X(k, i) */
            return gemm(t, tile(slot), RfpView{0, 1, leaf}, tile(slot), RfpView{0, 1, leaf}, m, m, m, 1.0, 0.0,
                        true, 0);
        const int64_t m1 = leaf_split(m, leaf), m2 = m - m1, s2 = slot + m1 / leaf;
        const RfpView t21 = t.sub(m1, 0);
        RFP_TRY(lauum(t, m1, slot));                                                                /* T11 := X11^T X11 */
        RFP_TRY(gemm(t, a, t21.t(), a, t21.t(), m1, m1, m2, 1.0, 1.0, true, 0));                   /* T11 += X21^T X21 */
        RFP_TRY(trmm_left(t21, m1, t.sub(m1, m1), m2, true, 1.0, s2));                             /* T21 := X22^T X21 */
        return lauum(t.sub(m1, m1), m2, s2);                                                       /* T22 := X22^T X22 */
    }
};

template <typename T>
static curfpStatus_t pftri_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, T *A)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if (!A) return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const int64_t leaf = rfp_leaf(Prec<T>::f64);
    const int64_t s1 = (g.n1 + leaf - 1) / leaf, s2 = (g.n2 + leaf - 1) / leaf;
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, (size_t)(s1 + s2) * leaf * leaf * sizeof(T), &ws);
    if (e != hipSuccess) return st_of(e);
    T *inv = (T *)ws;
    /* inverses of every diagonal leaf, both views in one launch */
    e = timed(h, CURFP_KERNEL_TRTRI, (double)n * leaf * leaf / 3.0, 0.0, [&] {
        return Prec<T>::trtri2(A, g.v11, g.n1, inv, g.v22, g.n2, inv + s1 * leaf * leaf, h->stream);
    });
    if (e != hipSuccess) return st_of(e);
    Inverter<T> iv{h, A, inv, leaf};
    const int64_t n1 = g.n1, n2 = g.n2;
    /* L^-1 over the views (the first split of the recursion at n1) */
    if ((e = iv.trtri(g.v11, n1, 0)) != hipSuccess) return st_of(e);
    if ((e = iv.trmm_right(g.v21, n2, g.v11, n1, -1.0, 0)) != hipSuccess) return st_of(e);
    if ((e = iv.trtri(g.v22, n2, s1)) != hipSuccess) return st_of(e);
    if ((e = iv.trmm_left(g.v21, n1, g.v22, n2, false, 1.0, s1)) != hipSuccess) return st_of(e);
    /* L^-T L^-1 */
    if ((e = iv.lauum(g.v11, n1, 0)) != hipSuccess) return st_of(e);
    if ((e = iv.gemm(g.v11, A, g.v21.t(), A, g.v21.t(), n1, n1, n2, 1.0, 1.0, true, 0)) != hipSuccess) return st_of(e);
    if ((e = iv.trmm_left(g.v21, n1, g.v22, n2, true, 1.0, s1)) != hipSuccess) return st_of(e);
    return st_of(iv.lauum(g.v22, n2, s1));
}

/* ------------------------------------------------------------------------ */
/* slansf / dlansf                                                           */
/* ------------------------------------------------------------------------ */

template <typename T>
static curfpStatus_t lansf_impl(curfpHandle_t h, curfpNormType_t norm, curfpOperation_t transr, curfpFillMode_t uplo,
                                int n, const T *arf, T *result)
{
    RFP_CHECK_HANDLE(h);
    if (!result || n < 0) return CURFP_STATUS_INVALID_VALUE;
    if ((norm != CURFP_NORM_MAX && norm != CURFP_NORM_ONE && norm != CURFP_NORM_FRO) ||
        (transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    *result = T(0);
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if (!arf) return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, lansf_scratch_bytes(g, (int)norm, sizeof(T)), &ws);
    if (e != hipSuccess) return st_of(e);
    const double nt = (double)n * (n + 1) / 2.0;
    e = timed(h, CURFP_KERNEL_OTHER, 0.0, nt * sizeof(T), [&] {
        if (sizeof(T) == 4) return launch_lansf_f32(g, (int)norm, (const float *)arf, ws, h->stream);
        return launch_lansf_f64(g, (int)norm, (const double *)arf, ws, h->stream);
    });
    unsigned long long acc[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(acc, ws, sizeof(acc), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return st_of(e);
    auto value_of = [](unsigned long long k) -> T {
        if (sizeof(T) == 4) {
            const uint32_t b = (uint32_t)k;
            float f;
            memcpy(&f, &b, 4);
            return (T)f;
        }
        double d;
        memcpy(&d, &k, 8);
        return (T)d;
    };
    if (norm == CURFP_NORM_MAX) {
        *result = value_of(acc[0]);
    } else if (norm == CURFP_NORM_ONE) {
        *result = value_of(acc[3]);
    } else {
        double all, dg;
        memcpy(&all, &acc[1], 8);
        memcpy(&dg, &acc[2], 8);
        *result = (T)sqrt(std::max(0.0, 2.0 * all - dg));
    }
    return CURFP_STATUS_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* ssfmm / dsfmm                                                             */
/* ------------------------------------------------------------------------ */

/* Reference src/curfp_ssfmm.cpp:107-306: side L  C := alpha A B + beta C (A m x m),
 * side R  C := alpha B A + beta C (A n x n); B and C column-major m x n.  Side R
 * runs as C^T = alpha A B^T + beta C^T on transposed views. */
template <typename T>
static curfpStatus_t sfmm_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, curfpSideMode_t side,
                               int m, int n, const T *alpha, const T *arf, const T *B, int ldb, const T *beta, T *C,
                               int ldc)
{
    RFP_CHECK_HANDLE(h);
    if (m < 0 || n < 0 || ldb < 1 || ldc < 1) return CURFP_STATUS_INVALID_VALUE;
    if (!alpha || !beta) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER) ||
        (side != CURFP_SIDE_LEFT && side != CURFP_SIDE_RIGHT))
        return CURFP_STATUS_INVALID_VALUE;
    if (m == 0 || n == 0) return CURFP_STATUS_SUCCESS;
    if (ldb < m || ldc < m || !B || !C || (*alpha != T(0) && !arf)) return CURFP_STATUS_INVALID_VALUE;
    const bool left = side == CURFP_SIDE_LEFT;
    const int64_t nA = left ? m : n, N = left ? n : m;
    const RfpGeom g = rfp_geom(nA, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const RfpView yv = left ? RfpView{0, 1, ldb} : RfpView{0, ldb, 1};
    const RfpView cv = left ? RfpView{0, 1, ldc} : RfpView{0, ldc, 1};
    const double fl = 2.0 * (double)nA * (double)nA * (double)N;
    const double by = (double)nA * (nA + 1) / 2.0 * sizeof(T) + 3.0 * (double)nA * N * sizeof(T);
    return st_of(timed(h, Prec<T>::f64 ? CURFP_KERNEL_GEMM_F64 : CURFP_KERNEL_GEMM_F32, fl, by, [&] {
        if (Prec<T>::f64)
            return launch_sfmm_f64((const double *)arf, g, (const double *)B, yv, (double *)C, cv, N,
                                   (double)*alpha, (double)*beta, h->stream);
        return launch_sfmm_f32((const float *)arf, g, (const float *)B, yv, (float *)C, cv, N, (float)*alpha,
                               (float)*beta, h->stream);
    }));
}

/* y := alpha A x + beta y with BLAS increments (element i at i * inc, or
 * (n - 1 - i) |inc| for inc < 0): each tile of A read once, its products
 * with x into per-tile slots, then one pass summing the slots. */
template <typename T>
static curfpStatus_t sfmv_impl(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const T *alpha,
                               const T *arf, const T *x, int incx, const T *beta, T *y, int incy)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || !alpha || !beta || incx == 0 || incy == 0) return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (n == 0) return CURFP_STATUS_SUCCESS;
    if (!y || (*alpha != T(0) && (!arf || !x))) return CURFP_STATUS_INVALID_VALUE;
    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const RfpView xv = vec_view(n, incx), yv = vec_view(n, incy);
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, symv_scratch_bytes(g, sizeof(T)), &ws);
    if (e != hipSuccess) return st_of(e);
    const double fl = 2.0 * (double)n * (double)n;
    const double by = (double)n * (n + 1) / 2.0 * sizeof(T) + 3.0 * (double)n * sizeof(T);
    return st_of(timed(h, CURFP_KERNEL_OTHER, fl, by, [&] {
        if (Prec<T>::f64)
            return launch_symv_f64(g, (const double *)arf, (const double *)x, xv, (double)*alpha, (double)*beta,
                                   (double *)y, yv, ws, h->stream);
        return launch_symv_f32(g, (const float *)arf, (const float *)x, xv, (float)*alpha, (float)*beta, (float *)y,
                               yv, ws, h->stream);
    }));
}

/* ------------------------------------------------------------------------ */
/* FP64 generator residual                                                   */
/* ------------------------------------------------------------------------ */

__global__ void residual_init_kernel(int64_t n, int64_t nrhs, double lambda, const double *X, int64_t ldx,
                                     const double *B, int64_t ldb, double *R, int64_t ldr)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx % n, c = idx / n;
    /* without B: R = lambda X, the first term of C X */
    R[i + c * ldr] = B ? B[i + c * ldb] - lambda * X[i + c * ldx] : lambda * X[i + c * ldx];
}

static int64_t residual_splits(int64_t n) { return n >= 8192 ? 32 : 1; }

static size_t residual_scratch_bytes(int64_t n, int64_t k, int64_t nrhs)
{
    return (size_t)k * nrhs * (1 + residual_splits(n)) * sizeof(double);
}

/* R = B - (alpha op(A) op(A)^T + lambda I) X in FP64, or R = C X when B is
 * null; `tmp` has residual_scratch_bytes of scratch. */
static hipError_t gen_residual_core(curfpHandle_t h, bool trans_n, int64_t n, int64_t k, int64_t nrhs, double alpha,
                                    const float *A, int64_t lda, double lambda, const double *X, int64_t ldx,
                                    const double *B, int64_t ldb, double *R, int64_t ldr, double *tmp)
{
    const int64_t total = n * nrhs;
    RFP_TRY(timed(h, CURFP_KERNEL_OTHER, 0.0, 24.0 * total, [&] {
        hipLaunchKernelGGL(residual_init_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, h->stream, n,
                           nrhs, lambda, X, ldx, B, ldb, R, ldr);
        return hipGetLastError();
    }));
    if (k == 0 || alpha == 0.0) return hipSuccess;
    /* T = op(A)^T X (k x nrhs), split-K over n; then R -= alpha op(A) T */
    const int64_t splits = residual_splits(n);
    const size_t t_elems = (size_t)k * nrhs;
    double *T = tmp;
    const RfpView av = trans_n ? RfpView{0, 1, lda} : RfpView{0, lda, 1};
    GemmArgs g1 = gemm_args();
    g1.a = A; g1.va = av.t();
    g1.b = X; g1.vb = RfpView{0, ldx, 1};
    g1.c = T; g1.vc = RfpView{0, 1, k};
    g1.M = k; g1.N = nrhs; g1.K = n;
    g1.alpha = 1.0; g1.beta = 0.0;
    g1.a_f32 = true;
    g1.splits = (int)splits;
    g1.partial = T + t_elems;
    RFP_TRY(run_gemm<double>(h, g1, RFP_MATH_F64));
    GemmArgs g2 = gemm_args();
    g2.a = A; g2.va = av;
    g2.b = T; g2.vb = RfpView{0, k, 1};
    g2.c = R; g2.vc = RfpView{0, 1, ldr};
    g2.M = n; g2.N = nrhs; g2.K = k;
    g2.alpha = B ? -alpha : alpha; g2.beta = 1.0;
    g2.a_f32 = true;
    return run_gemm<double>(h, g2, RFP_MATH_F64);
}

extern "C" curfpStatus_t curfpSgenResidual(curfpHandle_t h, curfpOperation_t trans, int n, int k, int nrhs,
                                           const double *alpha, const float *A, int lda, const double *lambda,
                                           const double *X, int ldx, const double *B, int ldb, double *R,
                                           int ldr)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || k < 0 || nrhs < 0 || lda < 1 || !alpha || !lambda) return CURFP_STATUS_INVALID_VALUE;
    if (trans != CURFP_OP_N && trans != CURFP_OP_T) return CURFP_STATUS_INVALID_VALUE;
    if (n == 0 || nrhs == 0) return CURFP_STATUS_SUCCESS;
    if (!X || !B || !R || ldx < n || ldb < n || ldr < n) return CURFP_STATUS_INVALID_VALUE;
    if (k > 0 && !A) return CURFP_STATUS_INVALID_VALUE;
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, residual_scratch_bytes(n, k, nrhs), &ws);
    if (e != hipSuccess) return st_of(e);
    return st_of(gen_residual_core(h, trans == CURFP_OP_N, n, k, nrhs, *alpha, A, lda, *lambda, X, ldx, B, ldb, R,
                                   ldr, (double *)ws));
}

/* ------------------------------------------------------------------------ */
/* Mixed-precision iterative refinement                                      */
/* ------------------------------------------------------------------------ */

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

/* Krylov steps per flexible-GMRES cycle of the refinement (fewer when the
 * basis of 2m + 1 blocks of n x nrhs FP64 would pass 8 GiB) */
constexpr int GMRES_RESTART = 20;

extern "C" curfpStatus_t curfpSpftrsRefine(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo,
                                           curfpOperation_t trans, int n, int k, int nrhs, const double *alpha,
                                           const float *A, int lda, const double *lambda, const float *F,
                                           const double *B, int ldb, double *X, int ldx, double tol, int max_iter,
                                           int *iters, double *resid)
{
    RFP_CHECK_HANDLE(h);
    if (n < 0 || k < 0 || nrhs < 0 || lda < 1 || !alpha || !lambda || max_iter < 1 || !(tol >= 0.0))
        return CURFP_STATUS_INVALID_VALUE;
    if ((transr != CURFP_OP_N && transr != CURFP_OP_T) || (trans != CURFP_OP_N && trans != CURFP_OP_T) ||
        (uplo != CURFP_FILL_MODE_LOWER && uplo != CURFP_FILL_MODE_UPPER))
        return CURFP_STATUS_INVALID_VALUE;
    if (iters) *iters = 0;
    if (resid) *resid = 0.0;
    if (n == 0 || nrhs == 0) return CURFP_STATUS_SUCCESS;
    if (!F || !B || !X || ldb < n || ldx < n || (k > 0 && !A)) return CURFP_STATUS_INVALID_VALUE;

    const RfpGeom g = rfp_geom(n, transr == CURFP_OP_N, uplo == CURFP_FILL_MODE_LOWER);
    const int64_t N = n, NR = nrhs, K = k;
    const int64_t gsplit = N >= 8192 ? 16 : 1;
    /* scratch shared by the residual and the Gram product, the solves'
     * (leaf inverses and sweep products, computed once and kept across the
     * solves), then the refinement's own buffers */
    size_t shared = std::max(residual_scratch_bytes(N, K, NR), (size_t)K * K * (1 + gsplit) * sizeof(double));
    shared = align256(shared);
    const size_t solve_bytes = align256(pftrs_scratch_bytes(g, false));
    const size_t r_bytes = align256((size_t)N * NR * sizeof(double));
    const size_t d_bytes = align256((size_t)N * NR * sizeof(float));
    void *ws = nullptr;
    hipError_t e = rfp_workspace(h, shared + solve_bytes + 2 * r_bytes + d_bytes + 256, &ws);
    if (e != hipSuccess) return st_of(e);
    char *base = (char *)ws;
    float *sv = (float *)(base + shared);
    double *R = (double *)(base + shared + solve_bytes);
    double *Xbest = (double *)(base + shared + solve_bytes + r_bytes);   /* the iterate of smallest residual */
    float *D = (float *)(base + shared + solve_bytes + 2 * r_bytes);
    double *sums = (double *)(base + shared + solve_bytes + 2 * r_bytes + d_bytes);  /* [0]=||R||^2 [1]=||X||^2 [2..3]=Gram */
    /* BF16 mode: the factor is BF16-accurate, so the sweeps' off-diagonal
     * products run in BF16 too; the refinement absorbs their rounding */
    const bool bf16_solve = math_of(h) == RFP_MATH_BF16;
    const double a = *alpha, lam = *lambda;
    const RfpView dv{0, 1, N};
    hipStream_t s = h->stream;

    /* ||C||_F^2 = a^2 ||A^T A||_F^2 + 2 a lam ||A||_F^2 + n lam^2, via the lower
     * triangle of G = op(A)^T op(A).  ||C|| only normalises the stopping test,
     * so G is formed in FP32 (relative error of order sqrt(n) u) on the FP32
     * MFMA path, split-K over n; the norms are summed in FP64. */
    double normC2 = (double)N * lam * lam;
    if (K > 0 && a != 0.0) {
        float *G = (float *)base;
        const RfpView av = trans == CURFP_OP_N ? RfpView{0, 1, lda} : RfpView{0, lda, 1};
        GemmArgs gg = gemm_args();
        gg.a = A; gg.va = av.t();
        gg.b = A; gg.vb = av.t();
        gg.c = G; gg.vc = RfpView{0, 1, K};
        gg.M = K; gg.N = K; gg.K = N;
        gg.alpha = 1.0; gg.beta = 0.0;
        gg.lower = true;
        gg.splits = (int)gsplit;
        gg.partial = G + K * K;
        if ((e = launch_zero_f64(sums, 4, s)) != hipSuccess) return st_of(e);
        if ((e = run_gemm<float>(h, gg, RFP_MATH_F32)) != hipSuccess) return st_of(e);
        if ((e = launch_gram_norms_f32(K, G, sums + 2, s)) != hipSuccess) return st_of(e);
        double hs[2];
        if ((e = hipMemcpyAsync(hs, sums + 2, sizeof(hs), hipMemcpyDeviceToHost, s)) != hipSuccess) return st_of(e);
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return st_of(e);
        normC2 += a * a * hs[0] + 2.0 * a * lam * hs[1];
    }
    const double normC = sqrt(normC2);

    /* X0 = F^{-1} B in FP32 */
    int it = 0;
    double rel = INFINITY;
    if ((e = launch_cast_f64_f32(N, NR, B, ldb, D, N, s)) != hipSuccess) return st_of(e);
    if ((e = pftrs_core<float>(h, g, F, D, dv, NR, sv, false, bf16_solve)) != hipSuccess) return st_of(e);
    if ((e = launch_acc_f32_f64(N, NR, D, N, X, ldx, false, s)) != hipSuccess) return st_of(e);
    it = 1;

    /* R = B - C X (FP64), ||R||_F, ||X||_F and the relative residual; the
     * read-back also collects a failed sweep's error word */
    double rn = INFINITY, nx = 0.0;
    auto residual = [&]() -> hipError_t {
        RFP_TRY(gen_residual_core(h, trans == CURFP_OP_N, N, K, NR, a, A, lda, lam, X, ldx, B, ldb, R, N,
                                  (double *)base));
        RFP_TRY(launch_zero_f64(sums, 2, s));
        RFP_TRY(launch_sumsq(N, NR, R, N, sums, s));
        RFP_TRY(launch_sumsq(N, NR, X, ldx, sums + 1, s));
        double hs[2] = {0.0, 0.0};
        int herr = 0;
        RFP_TRY(hipMemcpyAsync(hs, sums, sizeof(hs), hipMemcpyDeviceToHost, s));
        RFP_TRY(sweep_err_fetch(h, &herr));
        RFP_TRY(hipStreamSynchronize(s));
        RFP_TRY(sweep_err_check(h, herr));
        nx = sqrt(hs[1]);
        rn = sqrt(hs[0]);
        rel = (normC > 0.0 && nx > 0.0) ? rn / (normC * nx) : rn;
        return hipSuccess;
    };
    double best_rn = INFINITY, best_rel = INFINITY;
    bool x_is_best = false;
    auto track_best = [&]() {
        x_is_best = rn < best_rn;
        if (x_is_best) {
            best_rn = rn;
            best_rel = rel;
        }
    };

    /* Classical refinement: X += F^{-1} fp32(R) while each step shrinks the
     * residual at least fourfold.  A slower step means ||I - F^{-1} C|| is
     * not small (a BF16 factor against a condition number of ~10^3 already
     * diverges), and the correction equation C D = R is then solved by
     * flexible GMRES preconditioned with F instead. */
    bool krylov = false;
    double prev_rn = INFINITY;
    for (;;) {
        if ((e = residual()) != hipSuccess) return st_of(e);
        track_best();
        if (!(rel > tol) || it >= max_iter || !std::isfinite(rel)) break;
        if (it >= 2 && rn > 0.25 * prev_rn) {
            krylov = true;
            break;
        }
        prev_rn = rn;
        if (x_is_best && (e = launch_copy_f64(N, NR, X, ldx, Xbest, N, s)) != hipSuccess) return st_of(e);
        /* D = F^{-1} fp32(R); X += D */
        if ((e = launch_cast_f64_f32(N, NR, R, N, D, N, s)) != hipSuccess) return st_of(e);
        if ((e = pftrs_core<float>(h, g, F, D, dv, NR, sv, true, bf16_solve)) != hipSuccess) return st_of(e);
        if ((e = launch_acc_f32_f64(N, NR, D, N, X, ldx, true, s)) != hipSuccess) return st_of(e);
        ++it;
    }
    /* the last classical step made the residual worse: continue from (or
     * return) the best iterate */
    if (!x_is_best && std::isfinite(best_rn)) {
        if ((e = launch_copy_f64(N, NR, Xbest, N, X, ldx, s)) != hipSuccess) return st_of(e);
        if (krylov) {
            if ((e = residual()) != hipSuccess) return st_of(e);
            track_best();
        } else {
            rel = best_rel;
        }
    }

    if (krylov) {
        /*
         * Restarted flexible GMRES (Saad 1993) on C D = R, right-preconditioned
         * by the FP32 factor, one Krylov space per right-hand side — the
         * GMRES-IR of Carson & Higham (SIAM J. Sci. Comput. 40(2), 2018) with
         * the preconditioned directions Z_j = F^{-1} V_j kept.  Because C Z_j
         * is formed in FP64 from the generator and Z_j is stored exactly, the
         * Arnoldi relation C Z = V H holds to FP64 rounding whatever the
         * factor's accuracy, and the least-squares residual |g_{j+1}| of each
         * column is its true FP64 residual.  Orthogonalisation: classical
         * Gram-Schmidt applied twice.  Each step costs one FP32 solve (counted
         * in *iters), one FP64 product with C and one small read-back.
         */
        const size_t blk = (size_t)N * NR;
        int m = GMRES_RESTART;
        while (m > 4 && (size_t)(2 * m + 1) * blk * sizeof(double) > ((size_t)8 << 30)) --m;
        const int64_t nch = gm_chunks(N);
        const size_t hd_elems = (size_t)(2 * (m + 1) + 1) * NR;
        const size_t bytes = ((size_t)(2 * m + 1) * blk + (size_t)nch * (m + 1) * NR + hd_elems + (size_t)m * NR) *
                             sizeof(double);
        struct DevBuf {
            void *p = nullptr;
            ~DevBuf() { if (p) (void)hipFree(p); }
        } kb;
        if ((e = hipMalloc(&kb.p, bytes)) != hipSuccess) return st_of(e);
        double *V = (double *)kb.p;                       /* m + 1 basis blocks */
        double *Z = V + (size_t)(m + 1) * blk;            /* m preconditioned directions */
        double *part = Z + (size_t)m * blk;
        double *hd = part + (size_t)nch * (m + 1) * NR;   /* h (twice) and ||w||^2 of one step */
        double *yd = hd + hd_elems;
        double *h1 = hd, *h2 = hd + (size_t)(m + 1) * NR, *nrm2 = hd + (size_t)2 * (m + 1) * NR;
        std::vector<double> H((size_t)(m + 1) * m * NR), cs((size_t)m * NR), sn((size_t)m * NR),
            gv((size_t)(m + 1) * NR), y((size_t)m * NR), hh(hd_elems);
        auto Hc = [&](int64_t c, int i, int j) -> double & { return H[((size_t)c * m + j) * (m + 1) + i]; };
        double cycle_rn = rn;
        while (rel > tol && it < max_iter && std::isfinite(rel)) {
            /* V_0 = R / ||R|| per column */
            int herr = 0;
            if ((e = launch_gm_dots(N, NR, R, 1, R, part, nrm2, s)) != hipSuccess) return st_of(e);
            if ((e = launch_gm_normalize(N, NR, R, nrm2, V, s)) != hipSuccess) return st_of(e);
            if ((e = hipMemcpyAsync(hh.data() + 2 * (m + 1) * NR, nrm2, NR * sizeof(double), hipMemcpyDeviceToHost,
                                    s)) != hipSuccess)
                return st_of(e);
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return st_of(e);
            std::fill(gv.begin(), gv.end(), 0.0);
            for (int64_t c = 0; c < NR; ++c) gv[c] = sqrt(hh[2 * (m + 1) * NR + c]);
            /* stop the cycle once the columns' residuals meet the outer test with a margin */
            const double target = 0.5 * tol * normC * nx;
            int j = 0;
            while (j < m && it < max_iter) {
                double *Vj = V + (size_t)j * blk, *Zj = Z + (size_t)j * blk, *W = R;
                /* Z_j = F^{-1} V_j (FP32 solve), W = C Z_j (FP64) */
                if ((e = launch_cast_f64_f32(N, NR, Vj, N, D, N, s)) != hipSuccess) return st_of(e);
                if ((e = pftrs_core<float>(h, g, F, D, dv, NR, sv, true, bf16_solve)) != hipSuccess) return st_of(e);
                if ((e = launch_acc_f32_f64(N, NR, D, N, Zj, N, false, s)) != hipSuccess) return st_of(e);
                ++it;
                if ((e = gen_residual_core(h, trans == CURFP_OP_N, N, K, NR, a, A, lda, lam, Zj, N, nullptr, N, W, N,
                                           (double *)base)) != hipSuccess)
                    return st_of(e);
                /* W -= V h1; W -= V h2; V_{j+1} = W / ||W|| */
                if ((e = launch_gm_dots(N, NR, V, j + 1, W, part, h1, s)) != hipSuccess) return st_of(e);
                if ((e = launch_gm_update(N, NR, V, j + 1, h1, W, s)) != hipSuccess) return st_of(e);
                if ((e = launch_gm_dots(N, NR, V, j + 1, W, part, h2, s)) != hipSuccess) return st_of(e);
                if ((e = launch_gm_update(N, NR, V, j + 1, h2, W, s)) != hipSuccess) return st_of(e);
                if ((e = launch_gm_dots(N, NR, W, 1, W, part, nrm2, s)) != hipSuccess) return st_of(e);
                if ((e = launch_gm_normalize(N, NR, W, nrm2, V + (size_t)(j + 1) * blk, s)) != hipSuccess)
                    return st_of(e);
                if ((e = hipMemcpyAsync(hh.data(), hd, hd_elems * sizeof(double), hipMemcpyDeviceToHost, s)) !=
                    hipSuccess)
                    return st_of(e);
                if ((e = sweep_err_fetch(h, &herr)) != hipSuccess) return st_of(e);
                if ((e = hipStreamSynchronize(s)) != hipSuccess) return st_of(e);
                if ((e = sweep_err_check(h, herr)) != hipSuccess) return st_of(e);
                /* per column: new Hessenberg column, previous rotations, a new one */
                double est2 = 0.0;
                for (int64_t c = 0; c < NR; ++c) {
                    for (int i = 0; i <= j; ++i) Hc(c, i, j) = hh[(size_t)i * NR + c] + hh[(size_t)(m + 1 + i) * NR + c];
                    Hc(c, j + 1, j) = sqrt(hh[(size_t)2 * (m + 1) * NR + c]);
                    for (int i = 0; i < j; ++i) {
                        const double ci = cs[(size_t)i * NR + c], si = sn[(size_t)i * NR + c];
                        const double t = ci * Hc(c, i, j) + si * Hc(c, i + 1, j);
                        Hc(c, i + 1, j) = -si * Hc(c, i, j) + ci * Hc(c, i + 1, j);
                        Hc(c, i, j) = t;
                    }
                    const double p = Hc(c, j, j), q = Hc(c, j + 1, j), r = std::hypot(p, q);
                    const double cj = r > 0.0 ? p / r : 1.0, sj = r > 0.0 ? q / r : 0.0;
                    cs[(size_t)j * NR + c] = cj;
                    sn[(size_t)j * NR + c] = sj;
                    Hc(c, j, j) = r;
                    Hc(c, j + 1, j) = 0.0;
                    gv[(size_t)(j + 1) * NR + c] = -sj * gv[(size_t)j * NR + c];
                    gv[(size_t)j * NR + c] *= cj;
                    est2 += gv[(size_t)(j + 1) * NR + c] * gv[(size_t)(j + 1) * NR + c];
                }
                ++j;
                if (!(sqrt(est2) > target)) break;
            }
            /* y = H^{-1} g per column; X += Z y */
            for (int64_t c = 0; c < NR; ++c)
                for (int i = j - 1; i >= 0; --i) {
                    double t = gv[(size_t)i * NR + c];
                    for (int l = i + 1; l < j; ++l) t -= Hc(c, i, l) * y[(size_t)l * NR + c];
                    const double d = Hc(c, i, i);
                    y[(size_t)i * NR + c] = d != 0.0 ? t / d : 0.0;
                }
            if ((e = launch_copy_f64(N, NR, X, ldx, Xbest, N, s)) != hipSuccess) return st_of(e);
            if ((e = hipMemcpyAsync(yd, y.data(), (size_t)j * NR * sizeof(double), hipMemcpyHostToDevice, s)) !=
                hipSuccess)
                return st_of(e);
            if ((e = launch_gm_combine(N, NR, Z, j, yd, X, ldx, s)) != hipSuccess) return st_of(e);
            if ((e = residual()) != hipSuccess) return st_of(e);
            track_best();
            /* a cycle that did not reduce the residual (rounding floor reached): keep the previous iterate */
            if (!x_is_best) {
                if ((e = launch_copy_f64(N, NR, Xbest, N, X, ldx, s)) != hipSuccess) return st_of(e);
                rel = best_rel;
                break;
            }
            if (!(rn < 0.9 * cycle_rn)) break;
            cycle_rn = rn;
        }
        /* kb's buffer is freed after the work that reads it */
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return st_of(e);
    }
    if (iters) *iters = it;
    if (resid) *resid = rel;
    return CURFP_STATUS_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Strided GEMM                                                              */
/* ------------------------------------------------------------------------ */

template <typename T>
static curfpStatus_t gemm_strided_impl(curfpHandle_t h, int m, int n, int k, const T *alpha, const T *A,
                                       int64_t a_rs, int64_t a_cs, const T *B, int64_t b_rs, int64_t b_cs,
                                       const T *beta, T *C, int64_t c_rs, int64_t c_cs)
{
    RFP_CHECK_HANDLE(h);
    if (m < 0 || n < 0 || k < 0 || !alpha || !beta) return CURFP_STATUS_INVALID_VALUE;
    if (m == 0 || n == 0) return CURFP_STATUS_SUCCESS;
    if (!C || (k > 0 && *alpha != T(0) && (!A || !B))) return CURFP_STATUS_INVALID_VALUE;
    const bool unit_a = a_rs == 1 || a_cs == 1, unit_b = b_rs == 1 || b_cs == 1, unit_c = c_rs == 1 || c_cs == 1;
    if (!unit_c || (k > 0 && (!unit_a || !unit_b))) return CURFP_STATUS_INVALID_VALUE;
    GemmArgs g = gemm_args();
    g.a = A; g.va = RfpView{0, a_rs, a_cs};
    g.b = B; g.vb = RfpView{0, b_rs, b_cs};
    g.c = C; g.vc = RfpView{0, c_rs, c_cs};
    g.M = m; g.N = n; g.K = (*alpha == T(0)) ? 0 : k;
    g.alpha = (double)*alpha; g.beta = (double)*beta;
    return st_of(run_gemm<T>(h, g, Prec<T>::f64 ? RFP_MATH_F64 : RFP_MATH_F32));
}

extern "C" curfpStatus_t curfpSgemmStrided(curfpHandle_t h, int m, int n, int k, const float *alpha, const float *A,
                                           int64_t a_rs, int64_t a_cs, const float *B, int64_t b_rs, int64_t b_cs,
                                           const float *beta, float *C, int64_t c_rs, int64_t c_cs)
{ return gemm_strided_impl<float>(h, m, n, k, alpha, A, a_rs, a_cs, B, b_rs, b_cs, beta, C, c_rs, c_cs); }

extern "C" curfpStatus_t curfpDgemmStrided(curfpHandle_t h, int m, int n, int k, const double *alpha, const double *A,
                                           int64_t a_rs, int64_t a_cs, const double *B, int64_t b_rs, int64_t b_cs,
                                           const double *beta, double *C, int64_t c_rs, int64_t c_cs)
{ return gemm_strided_impl<double>(h, m, n, k, alpha, A, a_rs, a_cs, B, b_rs, b_cs, beta, C, c_rs, c_cs); }

/* ------------------------------------------------------------------------ */
/* C ABI                                                                     */
/* ------------------------------------------------------------------------ */

extern "C" {

curfpStatus_t curfpSsfrk(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, curfpOperation_t trans,
                         int n, int k, const float *alpha, const float *A, int lda, const float *beta, float *C)
{ return sfrk_impl<float>(h, transr, uplo, trans, n, k, alpha, A, lda, beta, C, false); }

curfpStatus_t curfpSsfrkBf16(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo,
                             curfpOperation_t trans, int n, int k, const float *alpha, const uint16_t *A,
                             int lda, const float *beta, float *C)
{ return sfrk_impl<float>(h, transr, uplo, trans, n, k, alpha, A, lda, beta, C, true); }

curfpStatus_t curfpDsfrk(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, curfpOperation_t trans,
                         int n, int k, const double *alpha, const double *A, int lda, const double *beta,
                         double *C)
{ return sfrk_impl<double>(h, transr, uplo, trans, n, k, alpha, A, lda, beta, C, false); }

curfpStatus_t curfpSpftrf(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, float *A,
                          int *info)
{ return pftrf_impl<float>(h, transr, uplo, n, A, info); }

curfpStatus_t curfpDpftrf(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, double *A,
                          int *info)
{ return pftrf_impl<double>(h, transr, uplo, n, A, info); }

curfpStatus_t curfpSpftrs(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, int nrhs,
                          const float *A, float *B, int ldb)
{ return pftrs_impl<float>(h, transr, uplo, n, nrhs, A, B, ldb, false); }

curfpStatus_t curfpSpftrsRowMajor(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                  int nrhs, const float *A, float *B, int ldb)
{ return pftrs_impl<float>(h, transr, uplo, n, nrhs, A, B, ldb, true); }

curfpStatus_t curfpDpftrs(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, int nrhs,
                          const double *A, double *B, int ldb)
{ return pftrs_impl<double>(h, transr, uplo, n, nrhs, A, B, ldb, false); }

curfpStatus_t curfpDpftrsRowMajor(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                  int nrhs, const double *A, double *B, int ldb)
{ return pftrs_impl<double>(h, transr, uplo, n, nrhs, A, B, ldb, true); }

curfpStatus_t curfpSstrttf(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                           const float *A, int lda, float *arf)
{ return trttf_impl<float>(h, transr, uplo, n, A, lda, arf); }

curfpStatus_t curfpDstrttf(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                           const double *A, int lda, double *arf)
{ return trttf_impl<double>(h, transr, uplo, n, A, lda, arf); }

curfpStatus_t curfpSstfttr(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                           const float *arf, float *A, int lda)
{ return tfttr_impl<float>(h, transr, uplo, n, arf, A, lda); }

curfpStatus_t curfpDstfttr(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                           const double *arf, double *A, int lda)
{ return tfttr_impl<double>(h, transr, uplo, n, arf, A, lda); }

curfpStatus_t curfpSaddToDiagonal(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                  const float *value, float *arf)
{ return add_diag_impl<float>(h, transr, uplo, n, value, arf); }

curfpStatus_t curfpDaddToDiagonal(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n,
                                  const double *value, double *arf)
{ return add_diag_impl<double>(h, transr, uplo, n, value, arf); }

curfpStatus_t curfpSsfr(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const float *alpha,
                        const float *x, int incx, float *arf)
{ return sfr_impl<float>(h, transr, uplo, n, alpha, x, incx, nullptr, 1, false, arf); }

curfpStatus_t curfpDsfr(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const double *alpha,
                        const double *x, int incx, double *arf)
{ return sfr_impl<double>(h, transr, uplo, n, alpha, x, incx, nullptr, 1, false, arf); }

curfpStatus_t curfpSsfr2(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const float *alpha,
                         const float *x, int incx, const float *y, int incy, float *arf)
{ return sfr_impl<float>(h, transr, uplo, n, alpha, x, incx, y, incy, true, arf); }

curfpStatus_t curfpDsfr2(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const double *alpha,
                         const double *x, int incx, const double *y, int incy, double *arf)
{ return sfr_impl<double>(h, transr, uplo, n, alpha, x, incx, y, incy, true, arf); }

curfpStatus_t curfpSsfr2k(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, curfpOperation_t trans, int n,
                          int k, const float *alpha, const float *A, int lda, const float *B, int ldb,
                          const float *beta, float *C)
{ return sfr2k_impl<float>(h, transr, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C); }

curfpStatus_t curfpDsfr2k(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, curfpOperation_t trans, int n,
                          int k, const double *alpha, const double *A, int lda, const double *B, int ldb,
                          const double *beta, double *C)
{ return sfr2k_impl<double>(h, transr, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C); }

curfpStatus_t curfpSsfmv(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const float *alpha,
                         const float *arf, const float *x, int incx, const float *beta, float *y, int incy)
{ return sfmv_impl<float>(h, transr, uplo, n, alpha, arf, x, incx, beta, y, incy); }

curfpStatus_t curfpDsfmv(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const double *alpha,
                         const double *arf, const double *x, int incx, const double *beta, double *y, int incy)
{ return sfmv_impl<double>(h, transr, uplo, n, alpha, arf, x, incx, beta, y, incy); }

curfpStatus_t curfpSpfcon(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const float *arf,
                          float anorm, float *rcond)
{ return pfcon_impl<float>(h, transr, uplo, n, arf, anorm, rcond); }

curfpStatus_t curfpDpfcon(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, const double *arf,
                          double anorm, double *rcond)
{ return pfcon_impl<double>(h, transr, uplo, n, arf, anorm, rcond); }

curfpStatus_t curfpSpftri(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, float *arf)
{ return pftri_impl<float>(h, transr, uplo, n, arf); }

curfpStatus_t curfpDpftri(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, int n, double *arf)
{ return pftri_impl<double>(h, transr, uplo, n, arf); }

curfpStatus_t curfpSsfmm(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, curfpSideMode_t side,
                         int m, int n, const float *alpha, const float *arf, const float *B, int ldb,
                         const float *beta, float *C, int ldc)
{ return sfmm_impl<float>(h, transr, uplo, side, m, n, alpha, arf, B, ldb, beta, C, ldc); }

curfpStatus_t curfpDsfmm(curfpHandle_t h, curfpOperation_t transr, curfpFillMode_t uplo, curfpSideMode_t side,
                         int m, int n, const double *alpha, const double *arf, const double *B, int ldb,
                         const double *beta, double *C, int ldc)
{ return sfmm_impl<double>(h, transr, uplo, side, m, n, alpha, arf, B, ldb, beta, C, ldc); }

curfpStatus_t curfpSlansf(curfpHandle_t h, curfpNormType_t norm, curfpOperation_t transr, curfpFillMode_t uplo,
                          int n, const float *arf, float *result)
{ return lansf_impl<float>(h, norm, transr, uplo, n, arf, result); }

curfpStatus_t curfpDlansf(curfpHandle_t h, curfpNormType_t norm, curfpOperation_t transr, curfpFillMode_t uplo,
                          int n, const double *arf, double *result)
{ return lansf_impl<double>(h, norm, transr, uplo, n, arf, result); }

}  /* extern "C" */
