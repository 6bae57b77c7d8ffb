/*
 * blas.hip — the plain GEMM updates on the vendor libraries.
 *
 * The reference hands its GEMM / SYRK updates to cuBLAS (e.g.
 * src/curfp_spftrf.cpp:270-274, src/curfp_ssfrk.cpp).  Here the rectangular
 * updates C := alpha A B^T + beta C whose operands are plain strided
 * matrices go to hipBLASLt (BF16 or FP32 operands, FP32 accumulation and
 * output) or rocBLAS (FP64), measured faster than the hand-written view GEMMs
 * on the chain's large shapes; everything with a fused epilogue (the BF16
 * shadow writes), the in-place solves, the operands converted while staged
 * and the mixed FP32 / FP64 residuals stay on the hand-written kernels, and
 * so do lower-triangular (SYRK) outputs, which the caller splits into
 * rectangles (for the library) and smaller triangles.
 *
 * Every view has a unit stride in one direction, so it is a column-major
 * matrix or the transpose of one:
 *   left operand  X (R x K):  rs == 1 -> stored R x K, op N;  cs == 1 -> stored K x R, op T
 *   right operand Y (C x K), used as Y^T:  cs == 1 -> stored K x C, op N;  rs == 1 -> stored C x K, op T
 *   output C (M x N):  rs == 1 -> C = op(A) op(B^T);  cs == 1 -> C^T = op(B) op(A^T)
 * Plans (descriptors, layouts and the heuristic's first algorithm) are cached
 * per shape on the handle.  rocBLAS runs with atomics off, so its results are
 * reproducible; the hipBLASLt algorithms chosen use no atomics either (checked
 * by the bit-identical repeat tests).
 */
#include "rfp_internal.h"
#include <hipblaslt/hipblaslt.h>
#include <rocblas/rocblas.h>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>

namespace {

constexpr size_t LT_WS_BYTES = size_t(64) << 20;   /* hipBLASLt workspace */

/* smallest dimension sent to the libraries (FP32 / BF16, FP64): below it the
 * hand-written kernels are as fast (measured on the chains: 256 instead of
 * 512 changes nothing for FP32 / BF16 and gains 3 ms on the FP64 chain) */
constexpr int64_t BLAS_MIN_DIM = 512, BLAS_MIN_DIM_F64 = 256;

struct Plan {
    bool ok = false;
    hipblasLtMatmulDesc_t op = nullptr;
    hipblasLtMatrixLayout_t l0 = nullptr, l1 = nullptr, lc = nullptr;
    hipblasLtMatmulAlgo_t algo{};
};

/* dtype code, op0, op1, m, n, k, ld0, ld1, ldc, beta == 0 */
typedef std::tuple<int, int, int, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int> Key;

struct BlasState {
    hipblasLtHandle_t lt = nullptr;
    rocblas_handle rb = nullptr;
    void *ws = nullptr;
    std::map<Key, Plan> plans;
};

/* one operand as a column-major matrix: pointer (element offset applied),
 * leading dimension, transposed or not */
struct Operand {
    int64_t off, ld;
    bool trans;
};

/* left operand X (R x K) of op(X) */
bool left_operand(const RfpView &v, int64_t R, int64_t K, Operand &o)
{
    if (v.rs == 1 && v.cs >= (R > 1 ? R : 1)) { o = Operand{v.off, v.cs, false}; return true; }
    if (v.cs == 1 && v.rs >= (K > 1 ? K : 1)) { o = Operand{v.off, v.rs, true}; return true; }
    return false;
}

/* right operand: Y (C x K) used as Y^T (K x C) */
bool right_operand(const RfpView &v, int64_t C, int64_t K, Operand &o)
{
    if (v.cs == 1 && v.rs >= (K > 1 ? K : 1)) { o = Operand{v.off, v.rs, false}; return true; }
    if (v.rs == 1 && v.cs >= (C > 1 ? C : 1)) { o = Operand{v.off, v.cs, true}; return true; }
    return false;
}

/* The product as D (m x n, column-major, ldc) = op(P0) op(P1) with P0 the
 * left operand of m rows and P1 the right operand of n rows. */
struct Problem {
    int64_t m, n, k, ldc, c_off;
    Operand p0, p1;
    bool swap;   /* C row-major: D = C^T, P0 = B, P1 = A */
};

bool map_problem(const GemmArgs &g, Problem &p)
{
    p.k = g.K;
    if (g.vc.rs == 1 && g.vc.cs >= (g.M > 1 ? g.M : 1)) {
        p.swap = false;
        p.m = g.M; p.n = g.N; p.ldc = g.vc.cs;
        if (!left_operand(g.va, g.M, g.K, p.p0) || !right_operand(g.vb, g.N, g.K, p.p1)) return false;
    } else if (g.vc.cs == 1 && g.vc.rs >= (g.N > 1 ? g.N : 1)) {
        p.swap = true;
        p.m = g.N; p.n = g.M; p.ldc = g.vc.rs;
        if (!left_operand(g.vb, g.N, g.K, p.p0) || !right_operand(g.va, g.M, g.K, p.p1)) return false;
    } else {
        return false;
    }
    p.c_off = g.vc.off;
    return true;
}

BlasState *state(curfpHandle_t h)
{
    if (h->blas) return static_cast<BlasState *>(h->blas);
    BlasState *s = new (std::nothrow) BlasState();
    if (!s) return nullptr;
    if (hipblasLtCreate(&s->lt) != HIPBLAS_STATUS_SUCCESS) s->lt = nullptr;
    if (rocblas_create_handle(&s->rb) != rocblas_status_success) {
        s->rb = nullptr;
    } else {
        (void)rocblas_set_atomics_mode(s->rb, rocblas_atomics_not_allowed);
        (void)rocblas_set_pointer_mode(s->rb, rocblas_pointer_mode_host);
    }
    if (hipMalloc(&s->ws, LT_WS_BYTES) != hipSuccess) s->ws = nullptr;
    h->blas = s;
    return s;
}

const Plan *lt_plan(BlasState *s, hipDataType in, const Problem &p, bool beta0)
{
    const Key key{(int)in, p.p0.trans, p.p1.trans, p.m, p.n, p.k, p.p0.ld, p.p1.ld, p.ldc, beta0};
    auto it = s->plans.find(key);
    if (it != s->plans.end()) return it->second.ok ? &it->second : nullptr;
    Plan &pl = s->plans[key];
    if (!s->lt || !s->ws) return nullptr;
    hipblasOperation_t t0 = p.p0.trans ? HIPBLAS_OP_T : HIPBLAS_OP_N, t1 = p.p1.trans ? HIPBLAS_OP_T : HIPBLAS_OP_N;
    bool ok = hipblasLtMatmulDescCreate(&pl.op, HIPBLAS_COMPUTE_32F, HIP_R_32F) == HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatmulDescSetAttribute(pl.op, HIPBLASLT_MATMUL_DESC_TRANSA, &t0, sizeof(t0)) == HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatmulDescSetAttribute(pl.op, HIPBLASLT_MATMUL_DESC_TRANSB, &t1, sizeof(t1)) == HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatrixLayoutCreate(&pl.l0, in, p.p0.trans ? p.k : p.m, p.p0.trans ? p.m : p.k, p.p0.ld) == HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatrixLayoutCreate(&pl.l1, in, p.p1.trans ? p.n : p.k, p.p1.trans ? p.k : p.n, p.p1.ld) == HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatrixLayoutCreate(&pl.lc, HIP_R_32F, p.m, p.n, p.ldc) == HIPBLAS_STATUS_SUCCESS;
    if (ok) {
        hipblasLtMatmulPreference_t pref = nullptr;
        size_t wsb = LT_WS_BYTES;
        hipblasLtMatmulHeuristicResult_t res{};
        int nres = 0;
        ok = hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS &&
             hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsb,
                                                   sizeof(wsb)) == HIPBLAS_STATUS_SUCCESS &&
             hipblasLtMatmulAlgoGetHeuristic(s->lt, pl.op, pl.l0, pl.l1, pl.lc, pl.lc, pref, 1, &res, &nres) ==
                 HIPBLAS_STATUS_SUCCESS &&
             nres > 0 && res.workspaceSize <= LT_WS_BYTES;
        if (pref) (void)hipblasLtMatmulPreferenceDestroy(pref);
        if (ok) pl.algo = res.algo;
    }
    pl.ok = ok;
    return ok ? &pl : nullptr;
}

}  // namespace

void blas_destroy(void *p)
{
    BlasState *s = static_cast<BlasState *>(p);
    if (!s) return;
    for (auto &kv : s->plans) {
        Plan &pl = kv.second;
        if (pl.lc) (void)hipblasLtMatrixLayoutDestroy(pl.lc);
        if (pl.l1) (void)hipblasLtMatrixLayoutDestroy(pl.l1);
        if (pl.l0) (void)hipblasLtMatrixLayoutDestroy(pl.l0);
        if (pl.op) (void)hipblasLtMatmulDescDestroy(pl.op);
    }
    if (s->lt) (void)hipblasLtDestroy(s->lt);
    if (s->rb) (void)rocblas_destroy_handle(s->rb);
    if (s->ws) (void)hipFree(s->ws);
    delete s;
}

bool blas_native_only()
{
    const char *e = std::getenv("CURFP_GEMM_NATIVE");
    return e && *e && std::strcmp(e, "0") != 0;
}

/* Which library (if any) takes this GEMM: 1 hipBLASLt BF16, 2 hipBLASLt FP32,
 * 3 rocBLAS FP64, 0 the hand-written kernels. */
int blas_route(const curfpHandle_t h, const GemmArgs &g, int math, bool f64)
{
    const RfpMath m = (RfpMath)math;
    if (h->blas_off || g.lower || g.inplace || g.c_sh_hi) return 0;
    const int64_t dmin = f64 ? BLAS_MIN_DIM_F64 : BLAS_MIN_DIM;
    if (g.M < dmin || g.N < dmin || g.K < dmin) return 0;
    Problem p;
    if (!map_problem(g, p)) return 0;
    if (f64) return (m == RFP_MATH_F64 && !g.a_f32 && !g.b_f32) ? 3 : 0;
    if (g.bf16_rows) {
        if (m != RFP_MATH_BF16 && m != RFP_MATH_BF16X3) return 0;
        if (m == RFP_MATH_BF16X3 && (!g.a_lo || !g.b_lo)) return 0;
        return 1;
    }
    if (g.a_bf16) return 1;
    return m == RFP_MATH_F32 ? 2 : 0;
}

hipError_t blas_gemm(curfpHandle_t h, const GemmArgs &g, int math, int route)
{
    const RfpMath m = (RfpMath)math;
    Problem p;
    if (!map_problem(g, p)) return hipErrorInvalidValue;
    BlasState *s = state(h);
    if (!s) return hipErrorOutOfMemory;
    const void *x = p.swap ? g.b : g.a, *y = p.swap ? g.a : g.b;
    if (route == 3) {
        if (!s->rb) return hipErrorNotSupported;
        if (rocblas_set_stream(s->rb, h->stream) != rocblas_status_success) return hipErrorUnknown;
        const double alpha = g.alpha, beta = g.beta;
        const double *P0 = (const double *)x + p.p0.off, *P1 = (const double *)y + p.p1.off;
        double *C = (double *)g.c + p.c_off;
        const rocblas_status st = rocblas_dgemm(
            s->rb, p.p0.trans ? rocblas_operation_transpose : rocblas_operation_none,
            p.p1.trans ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)p.m, (rocblas_int)p.n,
            (rocblas_int)p.k, &alpha, P0, (rocblas_int)p.p0.ld, P1, (rocblas_int)p.p1.ld, &beta, C,
            (rocblas_int)p.ldc);
        return st == rocblas_status_success ? hipSuccess : hipErrorUnknown;
    }
    const bool bf = route == 1;
    const Plan *pl = lt_plan(s, bf ? HIP_R_16BF : HIP_R_32F, p, g.beta == 0.0);
    if (!pl) return hipErrorNotSupported;
    const size_t es = bf ? 2 : 4;
    float *C = (float *)g.c + p.c_off;
    auto mm = [&](const void *a0, const void *a1, float alpha, float beta) {
        const char *P0 = (const char *)a0 + p.p0.off * (int64_t)es, *P1 = (const char *)a1 + p.p1.off * (int64_t)es;
        return hipblasLtMatmul(s->lt, pl->op, &alpha, P0, pl->l0, P1, pl->l1, &beta, C, pl->lc, C, pl->lc,
                               &pl->algo, s->ws, LT_WS_BYTES, h->stream) == HIPBLAS_STATUS_SUCCESS
                   ? hipSuccess
                   : hipErrorUnknown;
    };
    const float alpha = (float)g.alpha, beta = (float)g.beta;
    if (bf && g.bf16_rows && m == RFP_MATH_BF16X3) {
        /* split BF16: hi hi^T + hi lo^T + lo hi^T */
        const void *xl = p.swap ? g.b_lo : g.a_lo, *yl = p.swap ? g.a_lo : g.b_lo;
        hipError_t e = mm(x, y, alpha, beta);
        if (e == hipSuccess) e = mm(x, yl, alpha, 1.f);
        if (e == hipSuccess) e = mm(xl, y, alpha, 1.f);
        return e;
    }
    return mm(x, y, alpha, beta);
}
