/*
 * rfp_internal.h — handle layout, RFP geometry and launcher declarations
 * shared by the host drivers and the HIP kernels.
 *
 * RFP geometry as three strided views
 * -----------------------------------
 * LAPACK's RFP format (the reference's sub-block tables,
 * src/curfp_ssfrk.cpp:52-133, src/curfp_spftrf.cpp:62-180,
 * src/curfp_spftrs.cpp:60-213) splits the n x n symmetric matrix into a
 * leading n1 x n1 block, a trailing n2 x n2 block and the n2 x n1
 * off-diagonal block.  Here every one of the 8 storage variants is described
 * by the LOWER-triangular logical factorisation C = L*L^T (for uplo = U,
 * L = U^T), with each block of L (and of the unfactored C's lower triangle)
 * reached through a strided view
 *
 *     element (i, j) of the block  ->  arf[off + i*rs + j*cs]
 *
 * where (rs, cs) is (1, ld) ("column view") or (ld, 1) ("row view").  The
 * three views are
 *     v11: n1 x n1 lower triangle   (row view iff transr = T)
 *     v21: n2 x n1 rectangle        (column view iff (transr = N) == (uplo = L))
 *     v22: n2 x n2 lower triangle   (row view iff transr = N)
 * and the offsets/ld are LAPACK's.  This one description replaces the
 * reference's per-routine case tables: ssfrk writes the three blocks,
 * spftrf factors v11, solves v21, updates and factors v22, and spftrs runs
 * the forward and backward block substitutions on them.  The mapping is
 * checked element by element against LAPACK ?trttf for all 8 variants.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <vector>
#include "../../include/curfp.h"

struct RfpView {
    int64_t off, rs, cs;
    __host__ __device__ int64_t at(int64_t i, int64_t j) const { return off + i * rs + j * cs; }
    /* sub-view starting at block (i0, j0) */
    __host__ __device__ RfpView sub(int64_t i0, int64_t j0) const { return RfpView{at(i0, j0), rs, cs}; }
    /* transposed view: (i, j) of the result is (j, i) of this one */
    __host__ __device__ RfpView t() const { return RfpView{off, cs, rs}; }
};

struct RfpGeom {
    int64_t n, n1, n2, ld;
    RfpView v11, v21, v22;
};

/* LAPACK's n1/n2 split and sub-block offsets, expressed as the three views. */
static inline RfpGeom rfp_geom(int64_t n, bool transr_n, bool lower)
{
    RfpGeom g;
    g.n = n;
    int64_t o11, o21, o22;
    if (n % 2) {
        if (lower) { g.n1 = n - n / 2; g.n2 = n / 2; }
        else       { g.n1 = n / 2;     g.n2 = n - n / 2; }
        const int64_t n1 = g.n1, n2 = g.n2;
        if (transr_n) {
            g.ld = n;
            if (lower) { o11 = 0;  o21 = n1; o22 = n;  }
            else       { o11 = n2; o21 = 0;  o22 = n1; }
        } else if (lower) {
            g.ld = n1; o11 = 0; o21 = n1 * n1; o22 = 1;
        } else {
            g.ld = n2; o11 = n2 * n2; o21 = 0; o22 = n1 * n2;
        }
    } else {
        const int64_t k = n / 2;
        g.n1 = g.n2 = k;
        if (transr_n) {
            g.ld = n + 1;
            if (lower) { o11 = 1;     o21 = k + 1; o22 = 0; }
            else       { o11 = k + 1; o21 = 0;     o22 = k; }
        } else {
            g.ld = k;
            if (lower) { o11 = k;           o21 = k * (k + 1); o22 = 0;     }
            else       { o11 = k * (k + 1); o21 = 0;           o22 = k * k; }
        }
    }
    const bool row11 = !transr_n;
    const bool row22 = transr_n;
    const bool row21 = !(transr_n == lower);
    auto mk = [&](int64_t o, bool row) {
        return row ? RfpView{o, g.ld, 1} : RfpView{o, 1, g.ld};
    };
    g.v11 = mk(o11, row11);
    g.v21 = mk(o21, row21);
    g.v22 = mk(o22, row22);
    return g;
}

/*
 * Lower-triangular tile order used by every SYRK-shaped launch: tile-rows are
 * taken in bands of RFP_TILE_BAND; inside a band the tiles are visited column
 * by column, so the workgroups that run together on one XCD (consecutive ids
 * after the XCD remap) form a band-high block sharing a few row panels and a
 * few column panels in L2, instead of one row panel and many column panels.
 * Maps t in [0, T(T+1)/2) to (ti, tj), ti >= tj.
 */
#define RFP_TILE_BAND 8
__device__ __forceinline__ void tri_tile_coords(int64_t t, int64_t T, int64_t &ti, int64_t &tj)
{
    const int64_t G = RFP_TILE_BAND;
    int64_t b0 = 0;
    for (;;) {
        const int64_t g = (T - b0) < G ? (T - b0) : G;        /* rows in this band */
        const int64_t cnt = b0 * g + g * (g + 1) / 2;
        if (t < cnt || b0 + g >= T) {
            if (t < b0 * g) {                                    /* rectangular part */
                tj = t / g;
                ti = b0 + t % g;
            } else {                                             /* triangle part */
                int64_t u = t - b0 * g, c = 0;
                while (u >= g - c) { u -= g - c; ++c; }
                tj = b0 + c;
                ti = b0 + c + u;
            }
            return;
        }
        t -= cnt;
        b0 += g;
    }
}

/* ------------------------------------------------------------------------ */
/* Handle                                                                    */
/* ------------------------------------------------------------------------ */
/* One profiled launch: a pair of events on the handle stream around it. */
struct RfpProfRecord {
    int cls;
    double flops, bytes;
    hipEvent_t e0, e1;
};

struct RfpKernelStat {
    int64_t launches;
    double flops, bytes, ms;
};

struct curfpContext {
    hipStream_t     stream;
    int             device;
    curfpMathMode_t math;
    void           *ws;        /* cached device workspace (grown on demand) */
    size_t          ws_bytes;
    void           *sh;        /* cached BF16 shadow of the factor (BF16 math modes) */
    size_t          sh_bytes;
    int            *d_info;    /* device int: first failing minor (INT_MAX = none) */
    void           *sweep_sync;   /* SWEEP_SYNC_BYTES of sweep-solve progress words */
    unsigned long long sweep_epoch;
    int             sweep_wait_limit;  /* polls before a sweep workgroup's wait gives up */
    hipEvent_t      fence;     /* orders the previous stream's work before the next stream's (curfpSetStream) */
    /* kernel profiling (curfpSetProfiling / curfpGetKernelStats) */
    bool            prof;
    std::vector<RfpProfRecord> prof_pending;
    std::vector<hipEvent_t>    prof_pool;
    RfpKernelStat   prof_stats[CURFP_KERNEL_CLASSES];
    /* vendor-library GEMM state (blas.hip), created on first use */
    void           *blas;
    bool            blas_off;  /* CURFP_GEMM_NATIVE set: every GEMM on the hand-written kernels */
};

/*
 * Plain GEMM updates on the vendor libraries (blas.hip).  blas_route: 0 when
 * the hand-written kernels take g (lower-triangular outputs, in-place solves,
 * shadow epilogues, operands converted while staged, mixed precision, small
 * or non-unit-stride problems), else the library route for blas_gemm.
 */
int blas_route(const curfpHandle_t h, const struct GemmArgs &g, int math, bool f64);
hipError_t blas_gemm(curfpHandle_t h, const struct GemmArgs &g, int math, int route);
void blas_destroy(void *state);
bool blas_native_only();

/* Returns a device workspace of at least `bytes` bytes owned by the handle. */
hipError_t rfp_workspace(curfpHandle_t h, size_t bytes, void **out);

/* ------------------------------------------------------------------------ */
/* Kernel launchers (implemented in the .hip files)                          */
/* ------------------------------------------------------------------------ */

/* Compute precision of a view GEMM. */
enum RfpMath { RFP_MATH_F32 = 0, RFP_MATH_BF16 = 1, RFP_MATH_BF16X3 = 2, RFP_MATH_F64 = 3 };

/*
 * C(i, j) = alpha * sum_k A(i, k) * B(j, k) + beta * C(i, j)
 * for 0 <= i < M, 0 <= j < N (and only i >= j when `lower`, which requires
 * M == N and a square diagonal block).  A, B, C are strided views over the
 * given base pointers; each view must have a unit stride in one direction.
 * beta == 0 never reads C.  `inplace` = 1 / 2 declares that C aliases A
 * (resp. B); the kernel then gives one workgroup every column (resp. row)
 * of its output rows (columns), and the caller keeps N (resp. M) within one
 * tile (gemm_inplace_limit).
 */
struct GemmArgs {
    const void *a; RfpView va;
    const void *b; RfpView vb;
    void *c;       RfpView vc;
    int64_t M, N, K;
    double alpha, beta;
    bool lower;
    int  inplace;      /* 0 none, 1: C aliases A (one WG spans all N), 2: C aliases B (one WG spans all M) */
    bool a_bf16;       /* A and B are raw BF16 storage (ssfrk with BF16 input) */
    bool a_f32;        /* FP64 GEMM whose A is stored as FP32 (generator residual) */
    bool b_f32;        /* ... and whose B is stored as FP32 too (requires a_f32) */
    int  splits;       /* split-K slices (needs `partial`: splits*M*N elements) */
    void *partial;
    /* BF16 row-block operands (the factor's BF16 shadow): a / b are BF16 with
     * cs == 1; a_lo / b_lo are the split-BF16 low parts (BF16X3 only) */
    bool bf16_rows;
    const void *a_lo, *b_lo;
    /* FP32 GEMM epilogue also writes bf16(C) (and the low part) to these BF16
     * row blocks at view vcs — the shadow of a block that has become final */
    unsigned short *c_sh_hi, *c_sh_lo;
    RfpView vcs;
    /* batch > 1 (lower-triangular outputs, hand-written kernels only): the
     * same product on `batch` diagonal blocks, block z at element offsets
     * z * a_bs, z * b_bs, z * c_bs from block 0, the last block of order
     * m_last <= M */
    int batch;
    int64_t a_bs, b_bs, c_bs, m_last;
};
static inline GemmArgs gemm_args() { GemmArgs g{}; g.splits = 1; g.batch = 1; return g; }
hipError_t launch_gemm_f32(const GemmArgs &g, RfpMath math, hipStream_t s);
/* BF16 row-block GEMM with direct global -> LDS loads, bt x bt tiles (256 or
 * 128); returns hipErrorNotSupported when the operands do not meet its
 * requirements */
hipError_t launch_gemm_bf16_dma(const GemmArgs &g, int bt, hipStream_t s);
hipError_t launch_gemm_f64(const GemmArgs &g, hipStream_t s);
/* BF16 row blocks of an FP32 operand: hi[i * ldo + p] = bf16(A(i, p)) for
 * p < K, zero for K <= p < Kp (Kp <= ldo); lo (optional) the split-BF16 low
 * parts. */
hipError_t launch_to_bf16_rows(const float *A, RfpView av, int64_t rows, int64_t K, int64_t Kp, int64_t ldo,
                               unsigned short *hi, unsigned short *lo, hipStream_t s);

/* Largest N (inplace = 1) / M (inplace = 2) the in-place GEMM supports. */
int gemm_inplace_limit(bool f64);

/* Leaf size of the recursive factorisation / solves. */
int rfp_leaf(bool f64);

/* BF16 shadow geometry: L11 (n1 x n1, ld1), L21 (n2 x n1, ld1), L22 (n2 x n2, ld2) */
struct ShadowGeom {
    RfpView s11, s21, s22;
    int64_t elems;
};
inline ShadowGeom shadow_geom(const RfpGeom &g)
{
    const int64_t ld1 = (g.n1 + 7) / 8 * 8, ld2 = (g.n2 + 7) / 8 * 8;
    ShadowGeom sg;
    sg.s11 = RfpView{0, ld1, 1};
    sg.s21 = RfpView{g.n1 * ld1, ld1, 1};
    sg.s22 = RfpView{(g.n1 + g.n2) * ld1, ld2, 1};
    sg.elems = (g.n1 + g.n2) * ld1 + g.n2 * ld2;
    return sg;
}

/* Split point of the recursive factorisation and triangular solves: the
 * first ceil(blocks / 2) leaves of an m-row block. */
inline int64_t leaf_split(int64_t m, int64_t leaf)
{
    const int64_t nb = (m + leaf - 1) / leaf;
    return leaf * ((nb + 1) / 2);
}

/*
 * Diagonal tile Cholesky, in place on view v (m x m lower, m <= leaf).
 * On failure at local column j, atomicMin(info, info_base + j + 1).
 * When inv != nullptr, also writes inv(L) as a dense leaf x leaf row-major
 * lower-triangular matrix.  Skips all work if *info already holds a value
 * smaller than INT_MAX (an earlier block failed).
 */
hipError_t launch_potrf_tile_f32(float *base, RfpView v, int m, int *info, int info_base,
                                 float *inv, hipStream_t s);
hipError_t launch_potrf_tile_f64(double *base, RfpView v, int m, int *info, int info_base,
                                 double *inv, hipStream_t s);

/*
 * FP32 Cholesky of a diagonal block of leaf < m <= 2 leaf rows in one launch:
 * the two leaves, the panel below the first (also written to the BF16
 * shadow view sv when sh_hi != nullptr) and the update of the second (its
 * operands rounded to BF16 when bf16).  The two leaf inverses go to inv and
 * inv + leaf*leaf; info as launch_potrf_tile_f32 for the whole block.
 */
int potrf_pair_max();
hipError_t launch_potrf_pair_f32(float *base, RfpView v, int m, int *info, int info_base, float *inv,
                                 unsigned short *sh_hi, unsigned short *sh_lo, RfpView sv, bool bf16, hipStream_t s);

/*
 * Batched inverse of lower-triangular diagonal tiles: tile t covers rows
 * [t*leaf, min(m, (t+1)*leaf)) of the lower-triangular view v (m x m);
 * inv(tile t) goes to inv + t*leaf*leaf (row-major, zero upper part).
 */
hipError_t launch_trtri_tiles_f32(const float *base, RfpView v, int64_t m, float *inv, hipStream_t s);
hipError_t launch_trtri_tiles_f64(const double *base, RfpView v, int64_t m, double *inv, hipStream_t s);
/* The same for two views in one launch (their leaves' inverses at inv1, inv2) */
hipError_t launch_trtri_tiles2_f32(const float *base, RfpView v1, int64_t m1, float *inv1, RfpView v2, int64_t m2,
                                   float *inv2, hipStream_t s);
hipError_t launch_trtri_tiles2_f64(const double *base, RfpView v1, int64_t m1, double *inv1, RfpView v2, int64_t m2,
                                   double *inv2, hipStream_t s);

/*
 * FP32 block triangular solve (trsv_block.hip): B (m x nrhs view) :=
 * L^{-1} B (trans = false) or L^{-T} B (trans = true) for the m x m lower
 * block L (m <= block_solve_rows()), given the inverses of its 128-row leaves
 * at inv (leaf q at inv + q*128*128, row-major).
 */
int block_solve_rows();
hipError_t launch_block_solve_f32(const float *Lbase, RfpView Lv, int m, const float *inv, float *Bbase, RfpView Bv,
                                  int nrhs, bool trans, hipStream_t s);

/* The FP64 twin (m <= block_solve64_rows(), 64-row leaves whose inverses are
 * at inv + q*64*64, row-major, zero above the diagonal and beyond the leaf). */
int block_solve64_rows();
hipError_t launch_block_solve_f64(const double *Lbase, RfpView Lv, int m, const double *inv, double *Bbase,
                                  RfpView Bv, int nrhs, bool trans, hipStream_t s);

/*
 * FP32 X := X * L^{-T} (trsm_panel.hip) for X an mb x m view of the RFP array
 * and L an m x m lower-triangular view with m <= trsm_panel_max() (at most
 * four 128-row leaves, whose inverses are at inv + q*128*128, row-major).
 * When sh_hi is set the result is also rounded into the BF16 shadow view S
 * (and its low part into sh_lo).  bf16: the products with off-diagonal
 * leaves of L are BF16 products with FP32 accumulation (BF16 math mode).
 */
int trsm_panel_max();
hipError_t launch_trsm_panel_f32(float *a, RfpView X, int64_t mb, RfpView L, int64_t m, const float *inv,
                                 unsigned short *sh_hi, unsigned short *sh_lo, RfpView S, bool bf16, hipStream_t s);

/*
 * FP64 X := X * L^{-T} (trsm_panel64.hip) for X an mb x m view of the RFP
 * array and L an m x m lower-triangular view with m <= trsm_panel64_max() (at
 * most four 64-row leaves, whose inverses are at inv + q*64*64, row-major, zero
 * above the diagonal); both views need a unit stride.
 */
int trsm_panel64_max();
hipError_t launch_trsm_panel_f64(double *a, RfpView X, int64_t mb, RfpView L, int64_t m, const double *inv,
                                 hipStream_t s);

/*
 * FP32 substitution over the whole RFP factor in one launch (trsv_block.hip):
 * B (n x nrhs view) := L^{-1} B (trans = false) or L^{-T} B (trans = true),
 * given the inverses of the 128-row diagonal tiles of v11 then v22 at inv and
 * the products written by launch_sweep_pmat_f32 at pm.
 * `sync` holds SWEEP_SYNC_BYTES of device words owned by the handle (zeroed
 * once at creation); `epoch` must grow by one per launch on that handle.
 * A workgroup polls an earlier block's progress at most `wait_limit` times;
 * when that runs out it stops waiting, writes its block as NaN and sets the
 * error word (sweep_err_word), as does a workgroup whose ticket lies beyond
 * the grid (a counter shared by overlapping launches).
 * Supports nrhs <= 64 * SWEEP_MAX_GROUPS.  bf16: the off-diagonal
 * contributions (all but the last two blocks' of each row) are BF16 products
 * with FP32 accumulation — for solves inside iterative refinement of a
 * BF16-accurate factor; needs column-major B (bv.rs == 1).
 */
#define SWEEP_MAX_GROUPS 64
#define SWEEP_SYNC_BYTES 1024
hipError_t launch_sweep_solve_f32(const float *L, const RfpGeom &g, const float *inv, const float *pm, float *B,
                                  RfpView bv, int nrhs, bool trans, bool bf16, void *sync,
                                  unsigned long long epoch, int wait_limit, hipStream_t s);
/* the error word inside the SWEEP_SYNC_BYTES area (non-zero after a failed sweep) */
int *sweep_err_word(void *sync);
#define SWEEP_WAIT_LIMIT_DEFAULT (1 << 20)
/* C := alpha * A * Y + beta * C (trsv_block.hip) with A the n x n symmetric
 * matrix stored in RFP at A (geometry g), Y an n x N view and C an n x N view
 * (element (i, c) at off + i*rs + c*cs).  Y and C must not overlap.  FP32
 * on MFMA; FP64 on VALU FMAs. */
hipError_t launch_sfmm_f32(const float *A, const RfpGeom &g, const float *B, RfpView bv, float *C, RfpView cv,
                           int64_t N, float alpha, float beta, hipStream_t s);
hipError_t launch_sfmm_f64(const double *A, const RfpGeom &g, const double *B, RfpView bv, double *C, RfpView cv,
                           int64_t N, double alpha, double beta, hipStream_t s);
/* The sweep's products with the diagonal inverses, 4 * (blocks) 128 x 128
 * matrices at pm (written by this launch, read by both sweeps). */
hipError_t launch_sweep_pmat_f32(const float *L, const RfpGeom &g, const float *inv, float *pm, hipStream_t s);

/* RFP <-> dense row-major triangle copies (all 8 variants). */
hipError_t launch_trttf_f32(const RfpGeom &g, bool lower, const float *A, int64_t lda, float *arf, hipStream_t s);
hipError_t launch_trttf_f64(const RfpGeom &g, bool lower, const double *A, int64_t lda, double *arf, hipStream_t s);
hipError_t launch_tfttr_f32(const RfpGeom &g, bool lower, const float *arf, float *A, int64_t lda, hipStream_t s);
hipError_t launch_tfttr_f64(const RfpGeom &g, bool lower, const double *arf, double *A, int64_t lda, hipStream_t s);

/* arf[diag] += value; fill arf[0, nt) with zero. */
hipError_t launch_add_diag_f32(const RfpGeom &g, float value, float *arf, hipStream_t s);
hipError_t launch_add_diag_f64(const RfpGeom &g, double value, double *arf, hipStream_t s);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_set_int(int *p, int value, hipStream_t s);

/* Iterative-refinement helpers (refine.hip). */
hipError_t launch_cast_f64_f32(int64_t n, int64_t nrhs, const double *src, int64_t lds, float *dst, int64_t ldd,
                               hipStream_t s);
hipError_t launch_acc_f32_f64(int64_t n, int64_t nrhs, const float *D, int64_t ldd, double *X, int64_t ldx,
                              bool accumulate, hipStream_t s);
hipError_t launch_sumsq(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *out, hipStream_t s);
hipError_t launch_copy_f64(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *Y, int64_t ldy,
                           hipStream_t s);
/* Flexible GMRES on n x nrhs FP64 blocks, one Krylov space per column (basis
 * vector j at V + j*n*nrhs, leading dimension n):
 *   gm_dots:      out[i*nrhs + c] = <V_i(:, c), w(:, c)> for i < nv, through
 *                 `part` (gm_chunks(n) * nv * nrhs doubles), in a fixed order;
 *   gm_update:    w(:, c) -= sum_i h[i*nrhs + c] V_i(:, c);
 *   gm_normalize: v(:, c) = w(:, c) / sqrt(nrm2[c]);
 *   gm_combine:   X(:, c) += sum_j y[j*nrhs + c] Z_j(:, c). */
int64_t gm_chunks(int64_t n);
hipError_t launch_gm_dots(int64_t n, int64_t nrhs, const double *V, int nv, const double *w, double *part,
                          double *out, hipStream_t s);
hipError_t launch_gm_update(int64_t n, int64_t nrhs, const double *V, int nv, const double *h, double *w,
                            hipStream_t s);
hipError_t launch_gm_normalize(int64_t n, int64_t nrhs, const double *w, const double *nrm2, double *v, hipStream_t s);
hipError_t launch_gm_combine(int64_t n, int64_t nrhs, const double *Z, int nz, const double *y, double *X, int64_t ldx,
                             hipStream_t s);
/* out[0] += ||G||_F^2, out[1] += trace(G) for the symmetric k x k G given by
 * its lower triangle (column-major) */
hipError_t launch_gram_norms(int64_t k, const double *G, double *out, hipStream_t s);
hipError_t launch_gram_norms_f32(int64_t k, const float *G, double *out, hipStream_t s);
hipError_t launch_zero_f64(double *p, int n, hipStream_t s);

/*
 * Norms of the symmetric RFP matrix (lansf.hip): norm 0 = max |a_ij|,
 * 1 = largest absolute column sum, 2 = Frobenius.  ws holds
 * lansf_scratch_bytes; after the launches its first 32 bytes hold
 * [0] max |a| (bit pattern of T, u64), [1] sum of squares (double),
 * [2] sum of squared diagonal (double), [3] 1-norm (bit pattern of T, u64).
 */
size_t lansf_scratch_bytes(const RfpGeom &g, int norm, size_t elem);
hipError_t launch_lansf_f32(const RfpGeom &g, int norm, const float *arf, void *ws, hipStream_t s);
hipError_t launch_lansf_f64(const RfpGeom &g, int norm, const double *arf, void *ws, hipStream_t s);
/* y := alpha A x + beta y, A symmetric in RFP, x and y n-vectors as views
 * (element i at off + i * rs); ws holds symv_scratch_bytes (lansf.hip). */
size_t symv_scratch_bytes(const RfpGeom &g, size_t elem);
hipError_t launch_symv_f32(const RfpGeom &g, const float *arf, const float *x, RfpView xv, float alpha, float beta,
                           float *y, RfpView yv, void *ws, hipStream_t s);
hipError_t launch_symv_f64(const RfpGeom &g, const double *arf, const double *x, RfpView xv, double alpha,
                           double beta, double *y, RfpView yv, void *ws, hipStream_t s);

/*
 * Vector kernels of the condition estimate (pfcon.hip).  fill: mode 0 x = 1/n,
 * mode 1 x = e_j (0-based j), mode 2 x_i = (-1)^i (1 + i/(n-1)).  sign:
 * x_i = isgn_i = (x_i >= 0 ? 1 : -1).  stats: one workgroup writes *out
 * (isgn may be null: dot = 0; j < 0: xj = 0).
 */
struct PfconStats {
    double asum;   /* sum |x_i| */
    double dot;    /* sum isgn_i x_i */
    double xmax;   /* max |x_i| */
    double xj;     /* |x_j| */
    int64_t imax;  /* first index of max |x_i| */
    int64_t pad;
};
hipError_t launch_pfcon_fill_f32(float *x, int64_t n, int mode, int64_t j, hipStream_t s);
hipError_t launch_pfcon_fill_f64(double *x, int64_t n, int mode, int64_t j, hipStream_t s);
hipError_t launch_pfcon_sign_f32(float *x, float *isgn, int64_t n, hipStream_t s);
hipError_t launch_pfcon_sign_f64(double *x, double *isgn, int64_t n, hipStream_t s);
hipError_t launch_pfcon_stats_f32(const float *x, const float *isgn, int64_t n, int64_t j, PfconStats *out,
                                  hipStream_t s);
hipError_t launch_pfcon_stats_f64(const double *x, const double *isgn, int64_t n, int64_t j, PfconStats *out,
                                  hipStream_t s);
