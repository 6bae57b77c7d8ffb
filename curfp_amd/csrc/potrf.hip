/*
 * potrf.hip — diagonal-tile Cholesky and triangular inverse (one workgroup
 * per tile), the only non-MFMA arithmetic of spftrf / spftrs.
 *
 * The reference hands whole diagonal blocks to cusolverDnSpotrf
 * (src/curfp_spftrf.cpp:256, :278) and uses cublasStrsm for the solves.
 * Here the blocked recursion (rfp_drivers.hip) bottoms out in leaf tiles of
 * at most LEAF x LEAF; each leaf is factored in LDS by this kernel, which
 * also writes inv(L_leaf) so every triangular solve above the leaf level
 * becomes an MFMA GEMM.
 *
 * Failure detection follows LAPACK ?potrf: the first column j whose pivot is
 * not > 0 (or is NaN) reports info = j + 1 in the global numbering; the
 * atomicMin keeps the first failure, and later tiles skip their work once
 * info has been set.
 */
#include "rfp_internal.h"
#include <climits>
#include <type_traits>

template <typename T> struct LeafSize;
template <> struct LeafSize<float>  { static constexpr int value = 128; };
template <> struct LeafSize<double> { static constexpr int value = 64; };

int rfp_leaf(bool f64) { return f64 ? LeafSize<double>::value : LeafSize<float>::value; }

/* Load the lower triangle of an m x m view into s (pitch P), zero above;
 * all of a thread's loads are issued before its LDS stores. */
template <typename T, int NB, int P>
__device__ __forceinline__ void load_lower(T *s, const T *base, RfpView v, int m, int tid)
{
    constexpr int PER = NB * NB / 256;
    constexpr int STEP = 256 / NB;   /* rows (columns) between a thread's elements */
    T x[PER];
    if (v.rs == 1) {
        const int i = tid % NB, j0 = tid / NB;
        const T *p = base + v.off + i + (int64_t)j0 * v.cs;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int j = j0 + STEP * e;
            x[e] = (i < m && j < m && i >= j) ? p[(int64_t)STEP * e * v.cs] : T(0);
        }
#pragma unroll
        for (int e = 0; e < PER; ++e) s[i * P + j0 + STEP * e] = x[e];
    } else {
        const int j = tid % NB, i0 = tid / NB;
        const T *p = base + v.off + (int64_t)i0 * v.rs + j;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int i = i0 + STEP * e;
            x[e] = (i < m && j < m && i >= j) ? p[(int64_t)STEP * e * v.rs] : T(0);
        }
#pragma unroll
        for (int e = 0; e < PER; ++e) s[(i0 + STEP * e) * P + j] = x[e];
    }
}

template <typename T, int NB, int P>
__device__ __forceinline__ void store_lower(const T *s, T *base, RfpView v, int m, int tid)
{
    const int total = NB * NB;
    if (v.rs == 1) {
        for (int idx = tid; idx < total; idx += 256) {
            const int i = idx % NB, j = idx / NB;
            if (i < m && j < m && i >= j) base[v.off + i + (int64_t)j * v.cs] = s[i * P + j];
        }
    } else {
        for (int idx = tid; idx < total; idx += 256) {
            const int j = idx % NB, i = idx / NB;
            if (i < m && j < m && i >= j) base[v.off + (int64_t)i * v.rs + j * v.cs] = s[i * P + j];
        }
    }
}

/*
 * Register-blocked tile: a 16 x 16 grid of threads, thread (ty, tx) owning the
 * RB x RB block of rows ty*RB.., columns tx*RB.. (RB = NB / 16).  All
 * register arrays are indexed with compile-time indices only; runtime
 * positions (the current column j) are matched with selects.
 */
template <typename T, int NB>
struct RegTile {
    static constexpr int RB = NB / 16;
    int r0, c0, tx, ty;

    __device__ void init(int tid)
    {
        tx = tid & 15;
        ty = tid >> 4;
        r0 = ty * RB;
        c0 = tx * RB;
    }

    template <int P>
    __device__ void from_lds(const T *s, T (&a)[RB][RB]) const
    {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < RB; ++c) a[r][c] = s[(r0 + r) * P + c0 + c];
    }

    template <int P>
    __device__ void to_lds(T *s, const T (&a)[RB][RB]) const
    {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < RB; ++c) s[(r0 + r) * P + c0 + c] = a[r][c];
    }

    /* column jj (runtime, 0 <= jj < RB) of this thread's block */
    __device__ T pick_col(const T (&a)[RB][RB], int r, int jj) const
    {
        T v = 0;
#pragma unroll
        for (int c = 0; c < RB; ++c) v = (c == jj) ? a[r][c] : v;
        return v;
    }
};

/*
 * Right-looking Cholesky of the m x m lower tile held in registers.  Step j:
 * the threads owning column j publish it to LDS (double-buffered, so one
 * barrier per step), every thread reads the pivot and the 2*RB column
 * entries it needs, scales and applies the rank-1 update to its block.
 * Returns the failing column (pivot not > 0, or NaN), or -1.
 */
template <typename T, int NB>
__device__ int chol_regs(const RegTile<T, NB> &rt, T (&a)[NB / 16][NB / 16], T (*colbuf)[NB], int m)
{
    constexpr int RB = NB / 16;
    for (int j = 0; j < m; ++j) {
        T *buf = colbuf[j & 1];
        if (rt.tx == j / RB) {
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const int i = rt.r0 + r;
                if (i >= j) buf[i] = rt.pick_col(a, r, j % RB);
            }
        }
        __syncthreads();
        const T d = buf[j];
        if (!(d > T(0))) return j;          /* uniform: every thread read the same pivot */
        const T ljj = sqrt(d);
        const T rinv = T(1) / ljj;
        T li[RB], lk[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int i = rt.r0 + r;
            li[r] = (i > j && i < m) ? buf[i] * rinv : T(0);
        }
#pragma unroll
        for (int c = 0; c < RB; ++c) {
            const int k = rt.c0 + c;
            lk[c] = (k > j && k < m) ? buf[k] * rinv : T(0);
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int i = rt.r0 + r;
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                const int k = rt.c0 + c;
                if (k == j) {
                    if (i == j) a[r][c] = ljj;
                    else if (i > j) a[r][c] = li[r];
                } else if (k > j && i >= k) {
                    a[r][c] -= li[r] * lk[c];
                }
            }
        }
    }
    return -1;
}

/*
 * X = L^{-1} for the m x m lower-triangular L in registers `a`, by
 * right-looking forward elimination on X = I: step j publishes column j of L
 * and row j of X, then X(j, :) /= L(j, j) and X(i, :) -= L(i, j) X(j, :) for
 * i > j.  One barrier per step (double-buffered LDS rows / columns).
 */
template <typename T, int NB>
__device__ void inv_regs(const RegTile<T, NB> &rt, const T (&a)[NB / 16][NB / 16], T (&x)[NB / 16][NB / 16],
                         T (*colbuf)[NB], T (*rowbuf)[NB], int m)
{
    constexpr int RB = NB / 16;
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < RB; ++c) x[r][c] = (rt.r0 + r == rt.c0 + c && rt.r0 + r < m) ? T(1) : T(0);
    for (int j = 0; j < m; ++j) {
        T *cb = colbuf[j & 1];
        T *rb = rowbuf[j & 1];
        if (rt.tx == j / RB) {
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const int i = rt.r0 + r;
                if (i >= j) cb[i] = rt.pick_col(a, r, j % RB);
            }
        }
        if (rt.ty == j / RB) {
#pragma unroll
            for (int r = 0; r < RB; ++r)
                if (rt.r0 + r == j) {
#pragma unroll
                    for (int c = 0; c < RB; ++c) rb[rt.c0 + c] = x[r][c];
                }
        }
        __syncthreads();
        const T rinv = T(1) / cb[j];
        T xj[RB], lij[RB];
#pragma unroll
        for (int c = 0; c < RB; ++c) xj[c] = rb[rt.c0 + c] * rinv;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int i = rt.r0 + r;
            lij[r] = (i > j && i < m) ? cb[i] : T(0);
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int i = rt.r0 + r;
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                if (i == j) x[r][c] = xj[c];
                else if (i > j) x[r][c] -= lij[r] * xj[c];
            }
        }
    }
}

template <typename T, int NB, int P>
__device__ void write_inverse(const T *s, T *inv, int m, int tid)
{
    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int i = idx / NB, j = idx % NB;
        inv[idx] = (i < m && j <= i) ? s[i * P + j] : T(0);
    }
}

template <typename T, int NB>
__global__ __launch_bounds__(256) void potrf_tile_kernel(T *base, RfpView v, int m, int *info,
                                                         int info_base, T *inv)
{
    constexpr int P = NB + 1;
    constexpr int RB = NB / 16;
    __shared__ T s[NB * P];
    __shared__ T colbuf[2][NB];
    __shared__ T rowbuf[2][NB];
    const int tid = threadIdx.x;
    if (*info != INT_MAX) return;          /* an earlier block already failed */

    RegTile<T, NB> rt;
    rt.init(tid);
    load_lower<T, NB, P>(s, base, v, m, tid);
    __syncthreads();
    T a[RB][RB];
    rt.template from_lds<P>(s, a);
    const int fail = chol_regs<T, NB>(rt, a, colbuf, m);
    __syncthreads();
    rt.template to_lds<P>(s, a);
    __syncthreads();
    store_lower<T, NB, P>(s, base, v, m, tid);
    if (fail >= 0) {
        if (tid == 0) atomicMin(info, info_base + fail + 1);
        return;
    }
    if (inv) {
        T x[RB][RB];
        inv_regs<T, NB>(rt, a, x, colbuf, rowbuf, m);
        __syncthreads();
        rt.template to_lds<P>(s, x);
        __syncthreads();
        write_inverse<T, NB, P>(s, inv, m, tid);
    }
}

template <typename T, int NB>
__device__ __forceinline__ void trtri_tile(const T *base, RfpView v, int64_t m, T *inv, int64_t t)
{
    constexpr int P = NB + 1;
    constexpr int RB = NB / 16;
    __shared__ T s[NB * P];
    __shared__ T colbuf[2][NB];
    __shared__ T rowbuf[2][NB];
    const int tid = threadIdx.x;
    const int64_t r0 = t * NB;
    const int mt = (int)((m - r0) < NB ? (m - r0) : NB);
    RegTile<T, NB> rt;
    rt.init(tid);
    load_lower<T, NB, P>(s, base, v.sub(r0, r0), mt, tid);
    __syncthreads();
    T a[RB][RB], x[RB][RB];
    rt.template from_lds<P>(s, a);
    inv_regs<T, NB>(rt, a, x, colbuf, rowbuf, mt);
    __syncthreads();
    rt.template to_lds<P>(s, x);
    __syncthreads();
    write_inverse<T, NB, P>(s, inv + t * NB * NB, mt, tid);
}

/* Leaf inverses of up to two views in one launch: workgroups 0 .. nt1 - 1
 * invert view 1's leaves, the rest view 2's (two workgroups fit on a CU, so
 * both views' leaves run together instead of one launch after the other). */
template <typename T, int NB>
__global__ __launch_bounds__(256) void trtri_tiles_kernel(const T *base, RfpView v1, int64_t m1, T *inv1,
                                                          int64_t nt1, RfpView v2, int64_t m2, T *inv2)
{
    const int64_t t = blockIdx.x;
    if (t < nt1) trtri_tile<T, NB>(base, v1, m1, inv1, t);
    else         trtri_tile<T, NB>(base, v2, m2, inv2, t - nt1);
}

/* ------------------------------------------------------------------------ */
/* FP32 leaf: blocked, MFMA                                                  */
/* ------------------------------------------------------------------------ */
/*
 * The 128 x 128 FP32 leaf is factored as 4 x 4 blocks of 32 x 32 held in LDS:
 * for each block column p, one wave factors the diagonal block (one row per
 * lane, pivots and L(k, j) broadcast with readlane) and inverts it; the
 * panel below is solved as L(i, p) := L(i, p) * inv(L(p, p))^T and the
 * trailing blocks are updated, each as 32 x 32 x 32 v_mfma_f32_32x32x2_f32
 * products with the four waves taking different blocks.  The inverse of the
 * whole leaf is then assembled block by block,
 *     X(i, j) = -X(i, i) * sum_{j <= k < i} L(i, k) X(k, j),
 * again from MFMA block products.  A ragged leaf (m < 128) is padded with the
 * identity, which factors and inverts to itself.
 */
typedef float f32x16_t __attribute__((ext_vector_type(16)));

namespace leaf32 {

constexpr int NB = 128, BS = 32, NBLK = 4, LP = NB + 1;

/* acc += A * B over a 32 x 32 x 32 block product; element (i, k) of A at
 * A[i*ai + k*ak], element (k, j) of B at B[k*bk + j*bj] (all in LDS). */
__device__ __forceinline__ void mma(f32x16_t &acc, const float *A, int ai, int ak, const float *B, int bk, int bj,
                                    int lane)
{
    const int r = lane & 31, h = lane >> 5;
    /* all operands first: one LDS round trip for the 16 dependent MFMAs (the
     * scheduling barrier keeps the compiler from sinking each read to its
     * MFMA, which would put a round trip before every MFMA) */
    float a[BS / 2], b[BS / 2];
#pragma unroll
    for (int s = 0; s < BS / 2; ++s) {
        a[s] = A[r * ai + (2 * s + h) * ak];
        b[s] = B[(2 * s + h) * bk + r * bj];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < BS / 2; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[s], acc, 0, 0, 0);
}

__device__ __forceinline__ void zero(f32x16_t &acc)
{
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
}

/* C[i*ci + j] = alpha*acc (+ C when add) for the block held in acc */
__device__ __forceinline__ void store(const f32x16_t &acc, float *C, int ci, float alpha, bool add, int lane,
                                      bool lower_only = false)
{
    const int j = lane & 31;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (lower_only && i < j) continue;
        float *c = C + i * ci + j;
        *c = add ? *c + alpha * acc[r] : alpha * acc[r];
    }
}

__device__ __forceinline__ float bcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

/* order LDS traffic between the lanes of one wave */
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

/*
 * Factor the 32 x 32 diagonal block at D (pitch LP) with one wave, write L
 * into D's lower triangle and inv(L) into Xd's lower triangle.  Lane i < 32
 * holds row i of the block; lane 32 + c holds column c of X = inv(L),
 * starting from e_c.  W is a 16-byte aligned LDS scratch of 4 x BS floats
 * (the upper half is written by the inverse lanes and never read).
 * Returns the failing local column or -1 (wave-uniform).
 *
 * Right-looking: column j is published by the row lanes in LDS and read
 * back by every lane as 16-byte broadcasts c[k] = A(k, j).  Both kinds of
 * lane then apply the same step to their vector v:
 *     v[j] *= 1 / sqrt(A(j, j));   v[k] -= (v[j] / sqrt(A(j, j))) c[k],  k > j
 * which for a row lane is  L(i, j) = A(i, j) / L(j, j),
 * A(i, k) -= L(i, j) L(k, j), and for an inverse lane the column step of
 * forward substitution (x[j] /= L(j, j), x[k] -= L(k, j) x[j]), so one
 * instruction stream updates the factor rows and the inverse columns
 * together.  Software-pipelined: step j updates v[j + 1] first, publishes
 * it (column j + 1, final at that point) in the other half of W and issues
 * its reads before the rest of step j's updates, so the LDS round trip of
 * the next column overlaps the current one's arithmetic.  The update is
 * applied to all entries without a per-entry predicate: row entries right of
 * the diagonal (k > i) absorb updates but are never read (step j reads c[k]
 * only for k >= j, i.e. rows at or below the diagonal), so the block's upper
 * triangle in LDS holds garbage afterwards, which no later step reads; the
 * inverse's entries above the diagonal come out as exact zeros.
 * 1/sqrt(pivot) is v_rsq_f32 (1 ulp).
 */
__device__ int factor_diag(float *D, float *Xd, float *W, int lane)
{
    const int i = lane & 31;
    const bool row_lane = lane < 32;
    float v[BS];
    const float fi = (float)i;
#pragma unroll
    for (int k = 0; k < BS; ++k)
        v[k] = row_lane ? D[i * LP + k] : fmaxf(0.f, 1.f - fabsf(fi - (float)k));   /* row i, or e_i */
    /* column jn from W half h into c (entries jn .. BS-1, in 16-byte groups
     * held as vectors, so each group's LDS read lands in place) */
    typedef float f32x4c __attribute__((ext_vector_type(4)));
    auto fetch = [&](f32x4c (&c)[BS / 4], int jn, int h) {
        const f32x4c *w = reinterpret_cast<const f32x4c *>(W + h * BS);
#pragma unroll
        for (int g = jn / 4; g < BS / 4; ++g) c[g] = w[g];
    };
    f32x4c c[2][BS / 4];
    /* the inverse lanes' copies of the publication land in W[2 BS ..], which
     * nothing reads: no divergent branch splits the column loop */
    const int wpub = row_lane ? i : 2 * BS + i;
    W[wpub] = v[0];
    wave_sync();
    fetch(c[0], 0, 0);
    /* The pivot test is a select, not an exit, so the whole unrolled loop is
     * one basic block and the trailing updates of column j can be scheduled
     * into the LDS round trip of column j + 1.  After a failing pivot the
     * arithmetic runs on (NaN / garbage, never read back). */
    int fail = -1;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        const f32x4c (&cv)[BS / 4] = c[j & 1];
        auto cc = [&](int k) { return cv[k >> 2][k & 3]; };
        const float piv = cc(j);
        fail = (fail < 0 && !(piv > 0.f)) ? j : fail;
        const float rinv = __builtin_amdgcn_rsqf(piv);
        v[j] *= rinv;
        const float t = v[j] * rinv;
        if (j + 1 < BS) {
            v[j + 1] = fmaf(-t, cc(j + 1), v[j + 1]);
            /* half (j+1)&1 was last read at step j-1, whose reads this wave
             * has completed (they fed step j-1's updates) */
            W[((j + 1) & 1) * BS + wpub] = v[j + 1];
            wave_sync();
            fetch(c[(j + 1) & 1], j + 1, (j + 1) & 1);
        }
#pragma unroll
        for (int k = j + 2; k < BS; ++k) v[k] = fmaf(-t, cc(k), v[k]);
    }
    fail = __builtin_amdgcn_readfirstlane(fail);
    if (fail >= 0) return fail;
#pragma unroll
    for (int k = 0; k < BS; ++k) {
        if (row_lane) D[i * LP + k] = v[k];
        else          Xd[k * LP + i] = v[k];
    }
    return -1;
}

/* L(i, j) -= L(i, p) L(j, p)^T for one 32 x 32 block (lower part only when i == j) */
__device__ __forceinline__ void trail_block(float *L, int i, int j, int p, int lane)
{
    f32x16_t acc;
    zero(acc);
    mma(acc, L + (i * BS) * LP + p * BS, LP, 1, L + (j * BS) * LP + p * BS, 1, LP, lane);
    store(acc, L + (i * BS) * LP + j * BS, LP, -1.f, true, lane, i == j);
}

/* X(i, j) = -X(i, i) * sum_{k=j}^{i-1} L(i, k) X(k, j) for one block below
 * the diagonal of the inverse; Tw is the wave's own LDS scratch (pitch 33) */
__device__ __forceinline__ void inv_block(const float *L, float *X, float *Tw, int i, int j, int lane)
{
    f32x16_t acc;
    zero(acc);
    for (int k = j; k < i; ++k)
        mma(acc, L + (i * BS) * LP + k * BS, LP, 1, X + (k * BS) * LP + j * BS, LP, 1, lane);
    store(acc, Tw, 33, 1.f, false, lane);
    wave_sync();
    f32x16_t acc2;
    zero(acc2);
    mma(acc2, X + (i * BS) * LP + i * BS, LP, 1, Tw, 33, 1, lane);
    store(acc2, X + (i * BS) * LP + j * BS, LP, -1.f, false, lane);
    wave_sync();
}

/*
 * The leaf in phases, shared by the one-leaf kernel and the two-leaf kernel
 * below.  The factor lives in Lt and its inverse is built in Xt (both NB x LP
 * LDS arrays); Tt is the waves' 32 x 33 scratch.
 */

constexpr int PER = NB * NB / 256;   /* tile elements per thread */

/* Issue the loads of the lower triangle of the m x m tile (view v) into x.
 * Only the lower triangle is read by the factorisation; an entry above the
 * diagonal is loaded from the diagonal of its row instead (always inside the
 * view) and its value is never used.  Rows / columns at or beyond m are
 * clamped the same way (leaf_put replaces them by identity padding). */
__device__ __forceinline__ void leaf_fetch(float (&x)[PER], const float *base, RfpView v, int m, int tid)
{
    const int mc = m - 1;
    if (v.rs == 1) {
        /* column view: lanes run down a column */
        const int i = tid % NB, j0 = tid / NB, ic = min(i, mc);
        const float *p = base + v.off + ic;
#pragma unroll
        for (int e = 0; e < PER; ++e) x[e] = p[(int64_t)min(j0 + 2 * e, ic) * v.cs];
    } else {
        /* row view: lanes run along a row */
        const int j = tid % NB, i0 = tid / NB;
        const float *p = base + v.off;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int ic = min(i0 + 2 * e, mc);
            x[e] = p[(int64_t)ic * v.rs + min(j, ic) * v.cs];
        }
    }
}

/* The fetched tile into Lt, identity beyond m */
__device__ __forceinline__ void leaf_put(float *Lt, const float (&x)[PER], RfpView v, int m, int tid)
{
    if (v.rs == 1) {
        const int i = tid % NB, j0 = tid / NB;
#pragma unroll
        for (int e = 0; e < PER; ++e) Lt[i * LP + j0 + 2 * e] = x[e];
    } else {
        const int j = tid % NB, i0 = tid / NB;
#pragma unroll
        for (int e = 0; e < PER; ++e) Lt[(i0 + 2 * e) * LP + j] = x[e];
    }
    if (m < NB) {
        __syncthreads();
        for (int idx = tid; idx < NB * NB; idx += 256) {
            const int i = idx / NB, j = idx % NB;
            if (i >= m || j >= m) Lt[i * LP + j] = (i == j) ? 1.f : 0.f;
        }
    }
}

/* Load the m x m tile into Lt (all loads in flight before the LDS stores,
 * the check of info among them); returns false (storing nothing) when an
 * earlier tile has failed. */
__device__ __forceinline__ bool leaf_load(float *Lt, const float *base, RfpView v, int m, const int *info, int tid)
{
    float x[PER];
    leaf_fetch(x, base, v, m, tid);
    if (*info != INT_MAX) return false;
    leaf_put(Lt, x, v, m, tid);
    return true;
}

/*
 * Blocked right-looking over the 4 x 4 grid of 32 x 32 blocks.  Step p:
 *   A: wave 0 applies step p-1's update to the diagonal block (p, p) and
 *      factors it (with its inverse); meanwhile waves 1..3 apply step
 *      p-1's other trailing updates and assemble row p-1 of the inverse
 *      (its blocks need only earlier rows and X(p-1, p-1));
 *   B: panel  L(i, p) := L(i, p) X(p, p)^T  for the blocks below.
 * Returns the first failing local column or -1 (uniform); rows 0 .. 2 of
 * the inverse blocks are complete afterwards (leaf_inverse adds the last).
 */
__device__ __forceinline__ int leaf_factor(float *Lt, float *Xt, float (*Tt)[BS * 33], int *fail, bool want_inv,
                                           int tid)
{
    const int lane = tid & 63, wave = tid >> 6;
    if (tid == 0) *fail = -1;
    __syncthreads();
    for (int p = 0; p < NBLK; ++p) {
        if (wave == 0) {
            if (p > 0) {
                trail_block(Lt, p, p, p - 1, lane);
                wave_sync();
            }
            const int f = factor_diag(Lt + (p * BS) * LP + p * BS, Xt + (p * BS) * LP + p * BS, Tt[0], lane);
            if (lane == 0 && f >= 0) *fail = p * BS + f;
        } else if (p > 0) {
            int blk = 0;
            for (int i = p; i < NBLK; ++i)
                for (int j = p; j <= i; ++j) {
                    if (i == p && j == p) continue;
                    if (blk++ % 3 == wave - 1) trail_block(Lt, i, j, p - 1, lane);
                }
            if (want_inv)
                for (int j = wave - 1; j < p - 1; j += 3) inv_block(Lt, Xt, Tt[wave], p - 1, j, lane);
        }
        __syncthreads();
        if (*fail >= 0) break;
        const float *Xpp = Xt + (p * BS) * LP + p * BS;
        for (int i = p + 1 + wave; i < NBLK; i += 4) {
            float *Lip = Lt + (i * BS) * LP + p * BS;
            f32x16_t acc;
            zero(acc);
            mma(acc, Lip, LP, 1, Xpp, 1, LP, lane);   /* B(k, j) = X(j, k) */
            store(acc, Lip, LP, 1.f, false, lane);      /* the wave read all of Lip before writing */
        }
        __syncthreads();
    }
    return *fail;
}

/* The factor's lower triangle (m x m) back to the view */
__device__ __forceinline__ void leaf_store(const float *Lt, float *base, RfpView v, int m, int tid)
{
    /* a thread's 64 elements (every other column of its row, or every other
     * row of its column): all LDS reads, then the stores of those on or below
     * the diagonal and inside the tile */
    float x[PER];
    if (v.rs == 1) {
        const int i = tid % NB, j0 = tid / NB;
#pragma unroll
        for (int e = 0; e < PER; ++e) x[e] = Lt[i * LP + j0 + 2 * e];
        float *p = base + v.off + i + (int64_t)j0 * v.cs;
        const int64_t step = 2 * v.cs;
#pragma unroll
        for (int e = 0; e < PER; ++e)
            if (i < m && j0 + 2 * e <= i) p[e * step] = x[e];
    } else {
        const int j = tid % NB, i0 = tid / NB;
#pragma unroll
        for (int e = 0; e < PER; ++e) x[e] = Lt[(i0 + 2 * e) * LP + j];
        float *p = base + v.off + (int64_t)i0 * v.rs + j;
        const int64_t step = 2 * v.rs;
#pragma unroll
        for (int e = 0; e < PER; ++e)
            if (i0 + 2 * e < m && i0 + 2 * e >= j) p[e * step] = x[e];
    }
}

/* The last block row of the inverse, then the dense row-major inverse
 * (zero above the diagonal and beyond m) to inv; Xt holds the whole inverse
 * afterwards (its blocks above the diagonal blocks are not written). */
__device__ __forceinline__ void leaf_inverse(const float *Lt, float *Xt, float (*Tt)[BS * 33], float *inv, int m,
                                             int tid)
{
    const int lane = tid & 63, wave = tid >> 6;
    if (wave < NBLK - 1) inv_block(Lt, Xt, Tt[wave], NBLK - 1, wave, lane);
    __syncthreads();
    /* each thread's 16 pieces of 4: all LDS reads before the stores */
    float w[NB * NB / 1024][4];
#pragma unroll
    for (int e = 0; e < NB * NB / 1024; ++e) {
        const int idx = 4 * (tid + 256 * e);
#pragma unroll
        for (int u = 0; u < 4; ++u) w[e][u] = Xt[(idx / NB) * LP + idx % NB + u];
    }
#pragma unroll
    for (int e = 0; e < NB * NB / 1024; ++e) {
        const int idx = 4 * (tid + 256 * e);
        const int i = idx / NB, j = idx % NB;
        float o[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) o[u] = (i < m && j + u <= i) ? w[e][u] : 0.f;
        *reinterpret_cast<float4 *>(inv + idx) = make_float4(o[0], o[1], o[2], o[3]);
    }
}

__global__ __launch_bounds__(256) void potrf_leaf_kernel(float *base, RfpView v, int m, int *info, int info_base,
                                                         float *inv)
{
    __shared__ float L[NB * LP];
    __shared__ float X[NB * LP];
    __shared__ __attribute__((aligned(16))) float T[4][BS * 33];
    __shared__ int fail;
    const int tid = threadIdx.x;
    if (!leaf_load(L, base, v, m, info, tid)) return;
    const int f = leaf_factor(L, X, T, &fail, inv != nullptr, tid);
    /* write L back (also after a failure: LAPACK leaves a partial factor);
     * the stores drain while the inverse is assembled */
    leaf_store(L, base, v, m, tid);
    if (f >= 0) {
        if (tid == 0) atomicMin(info, info_base + f + 1);
        return;
    }
    if (inv) leaf_inverse(L, X, T, inv, m, tid);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

/* acc += A * B^T for one 32 x 32 x 32 block product with A(i, k) at
 * A[i*LP + k] and B(j, k) at B[j*LP + k] (LDS), both operands rounded to
 * BF16 (nearest even): two v_mfma_f32_32x32x16_bf16, exact products with
 * FP32 accumulation (the arithmetic of the BF16 updates).  Lane l supplies
 * row l % 32 of each operand at k = 16 t + 8 (l / 32) + 0 .. 7; the result
 * has mma()'s layout. */
__device__ __forceinline__ void mma_bf16t(f32x16_t &acc, const float *A, const float *B, int lane)
{
    const int r = lane & 31, h = lane >> 5;
    float a[2][8], b[2][8];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[t][e] = A[r * LP + 16 * t + 8 * h + e];
            b[t][e] = B[r * LP + 16 * t + 8 * h + e];
        }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        bf16x8_t av, bv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            av[e] = (__bf16)a[t][e];
            bv[e] = (__bf16)b[t][e];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
}

/* mma_bf16t with both operands split into BF16 high and low parts (three
 * MFMAs per 16-wide k step: hi hi^T + hi lo^T + lo hi^T, products to ~2^-16
 * relative): the split-BF16 inverse products of the BF16-mode panels. */
__device__ __forceinline__ void mma_bf16x3t(f32x16_t &acc, const float *A, const float *B, int lane)
{
    const int r = lane & 31, h = lane >> 5;
    float a[2][8], b[2][8];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[t][e] = A[r * LP + 16 * t + 8 * h + e];
            b[t][e] = B[r * LP + 16 * t + 8 * h + e];
        }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        bf16x8_t ah, al, bh, bl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            ah[e] = (__bf16)a[t][e];
            al[e] = (__bf16)(a[t][e] - (float)ah[e]);
            bh[e] = (__bf16)b[t][e];
            bl[e] = (__bf16)(b[t][e] - (float)bh[e]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    }
}

/*
 * Two leaves of a 128 < m <= 256 diagonal block in one launch: the
 * recursion's potrf(m) = potrf(L00); X := C10 L00^-T; C11 -= X X^T;
 * potrf(C11), which otherwise takes a leaf launch, a panel launch, an
 * update launch and a second leaf launch (the bottom of spftrf's recursion).
 * One workgroup: leaf 0 is factored as in potrf_leaf_kernel (its inverse
 * written to inv); C10 (the m - 128 rows below it) is staged into L's LDS
 * and multiplied by the inverse still in X (the panel's rows, written to the
 * view and, in the BF16 modes, to the BF16 shadow view sv); C11 is loaded
 * into X, updated with the panel's product (BF16-rounded operands when bf16,
 * like the BF16 updates; the panel itself split-BF16 then, like the BF16-mode
 * panel TRSM), and factored as leaf 1 with the roles of the two
 * LDS arrays exchanged.  info as LAPACK's (first failing column of the
 * block).
 */
__global__ __launch_bounds__(256) void potrf_pair_kernel(float *base, RfpView v, int m, int *info, int info_base,
                                                         float *inv, unsigned short *sh_hi, unsigned short *sh_lo,
                                                         RfpView sv, int bf16)
{
    __shared__ float L[NB * LP];
    __shared__ float X[NB * LP];
    __shared__ __attribute__((aligned(16))) float T[4][BS * 33];
    __shared__ int fail;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m2 = m - NB;   /* rows of leaf 1 (1 .. NB) */

    /* leaf 0 */
    if (!leaf_load(L, base, v, NB, info, tid)) return;
    const int f0 = leaf_factor(L, X, T, &fail, true, tid);
    /* the loads of C10 (the m2 x 128 rows below leaf 0; rows at or beyond m2
     * clamped, zeroed in LDS) are in flight during leaf 0's store and inverse */
    const RfpView v10 = v.sub(NB, 0);
    float x10[PER];
    if (v10.rs == 1) {
        const float *p = base + v10.off + min(tid % NB, m2 - 1);
#pragma unroll
        for (int e = 0; e < PER; ++e) x10[e] = p[(int64_t)(tid / NB + 2 * e) * v10.cs];
    } else {
        const float *p = base + v10.off + tid % NB;
#pragma unroll
        for (int e = 0; e < PER; ++e) x10[e] = p[(int64_t)min(tid / NB + 2 * e, m2 - 1) * v10.rs];
    }
    leaf_store(L, base, v, NB, tid);
    if (f0 >= 0) {
        if (tid == 0) atomicMin(info, info_base + f0 + 1);
        return;
    }
    leaf_inverse(L, X, T, inv, NB, tid);
    __syncthreads();   /* L's store and every read of it done */
    if (v10.rs == 1) {
        const int i = tid % NB, j0 = tid / NB;
#pragma unroll
        for (int e = 0; e < PER; ++e) L[i * LP + j0 + 2 * e] = i < m2 ? x10[e] : 0.f;
    } else {
        const int j = tid % NB, i0 = tid / NB;
#pragma unroll
        for (int e = 0; e < PER; ++e) L[(i0 + 2 * e) * LP + j] = i0 + 2 * e < m2 ? x10[e] : 0.f;
    }
    /* C11's loads, in flight during the panel */
    const RfpView v11 = v.sub(NB, NB);
    float x11[PER];
    leaf_fetch(x11, base, v11, m2, tid);
    __syncthreads();

    /* panel: wave w computes its row slab P(w, jb) = sum_{kb <= jb} C10(w, kb) inv(jb, kb)^T,
     * jb = 0 .. 3 (ten block products per wave), and overwrites C10(w, .) with it */
    {
        f32x16_t acc[NBLK];
#pragma unroll
        for (int jb = 0; jb < NBLK; ++jb) {
            zero(acc[jb]);
#pragma unroll
            for (int kb = 0; kb <= jb; ++kb) {
                const float *Ak = L + (wave * BS) * LP + kb * BS, *Bk = X + (jb * BS) * LP + kb * BS;
                if (bf16) mma_bf16x3t(acc[jb], Ak, Bk, lane);   /* as the BF16-mode panel TRSM */
                else      mma(acc[jb], Ak, LP, 1, Bk, 1, LP, lane);
            }
        }
        wave_sync();   /* the slab is the wave's own: its reads are done before it is overwritten */
#pragma unroll
        for (int jb = 0; jb < NBLK; ++jb) store(acc[jb], L + (wave * BS) * LP + jb * BS, LP, 1.f, false, lane);
    }
    __syncthreads();
    /* the panel to the view and the BF16 shadow; C11 (its lower triangle,
     * identity beyond m2) into X, whose inverse of leaf 0 is written out */
    {
        /* all LDS reads of the thread's 64 elements, then the stores of the
         * rows inside the panel */
        float x[PER];
        if (v10.rs == 1) {
            const int i = tid % NB, j0 = tid / NB;
#pragma unroll
            for (int e = 0; e < PER; ++e) x[e] = L[i * LP + j0 + 2 * e];
            float *p = base + v10.off + i + (int64_t)j0 * v10.cs;
#pragma unroll
            for (int e = 0; e < PER; ++e)
                if (i < m2) p[e * 2 * v10.cs] = x[e];
        } else {
            const int j = tid % NB, i0 = tid / NB;
#pragma unroll
            for (int e = 0; e < PER; ++e) x[e] = L[(i0 + 2 * e) * LP + j];
            float *p = base + v10.off + (int64_t)i0 * v10.rs + j;
#pragma unroll
            for (int e = 0; e < PER; ++e)
                if (i0 + 2 * e < m2) p[e * 2 * v10.rs] = x[e];
        }
    }
    if (sh_hi) {
        const RfpView s10 = sv.sub(NB, 0);
        if (s10.cs == 1 && s10.rs % 8 == 0 && s10.off % 8 == 0) {
            /* rows k-contiguous and 16-byte aligned: 8 elements per store */
            constexpr int CH = NB / 8;
#pragma unroll
            for (int e = 0; e < NB * CH / 256; ++e) {
                const int idx = tid + 256 * e, i = idx / CH, c = idx % CH;
                if (i >= m2) continue;
                bf16x8_t hv, lv;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float x = L[i * LP + 8 * c + q];
                    hv[q] = (__bf16)x;   /* nearest even */
                    lv[q] = (__bf16)(x - (float)hv[q]);
                }
                const int64_t so = s10.off + (int64_t)i * s10.rs + 8 * c;
                *reinterpret_cast<bf16x8_t *>(sh_hi + so) = hv;
                if (sh_lo) *reinterpret_cast<bf16x8_t *>(sh_lo + so) = lv;
            }
        } else {
            for (int idx = tid; idx < NB * NB; idx += 256) {
                const int i = s10.cs == 1 ? idx / NB : idx % NB, j = s10.cs == 1 ? idx % NB : idx / NB;
                if (i >= m2) continue;
                const float x = L[i * LP + j];
                const __bf16 hb = (__bf16)x;
                const int64_t so = s10.off + (int64_t)i * s10.rs + (int64_t)j * s10.cs;
                sh_hi[so] = __builtin_bit_cast(unsigned short, hb);
                if (sh_lo) sh_lo[so] = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)hb));
            }
        }
    }
    leaf_put(X, x11, v11, m2, tid);
    __syncthreads();

    /* C11 -= P P^T on the 10 blocks on or below the diagonal, block t =
     * (ib, jb) in row order.  Wave w takes blocks 2w and 2w + 1 whole and half
     * the k range of block 8 + (w & 1) (waves 0, 1 the first half, stored
     * first; waves 2, 3 the second, stored after a barrier), ten block
     * products each. */
    {
        auto blk_i = [](int t) { return (int)((0x3333222110ull >> (4 * t)) & 15); };
        auto blk_j = [](int t) { return (int)((0x3210210100ull >> (4 * t)) & 15); };
        auto syrk = [&](f32x16_t &c, int t, int k0, auto nk) {
            const int ib = blk_i(t), jb = blk_j(t);
            zero(c);
#pragma unroll
            for (int q = 0; q < decltype(nk)::value; ++q) {
                const int kb = k0 + q;
                const float *Ai = L + (ib * BS) * LP + kb * BS, *Bj = L + (jb * BS) * LP + kb * BS;
                if (bf16) mma_bf16t(c, Ai, Bj, lane);
                else      mma(c, Ai, LP, 1, Bj, 1, LP, lane);
            }
        };
        auto put = [&](const f32x16_t &c, int t) {
            const int ib = blk_i(t), jb = blk_j(t);
            store(c, X + (ib * BS) * LP + jb * BS, LP, -1.f, true, lane, ib == jb);
        };
        f32x16_t c0, c1, ch;
        const int th = 8 + (wave & 1), k0 = wave < 2 ? 0 : NBLK / 2;
        syrk(c0, 2 * wave, 0, std::integral_constant<int, NBLK>{});
        syrk(c1, 2 * wave + 1, 0, std::integral_constant<int, NBLK>{});
        syrk(ch, th, k0, std::integral_constant<int, NBLK / 2>{});
        put(c0, 2 * wave);
        put(c1, 2 * wave + 1);
        if (wave < 2) put(ch, th);
        __syncthreads();
        if (wave >= 2) put(ch, th);
    }
    __syncthreads();

    /* leaf 1: factor X in place, its inverse in L */
    const int f1 = leaf_factor(X, L, T, &fail, true, tid);
    leaf_store(X, base, v.sub(NB, NB), m2, tid);
    if (f1 >= 0) {
        if (tid == 0) atomicMin(info, info_base + NB + f1 + 1);
        return;
    }
    leaf_inverse(X, L, T, inv + NB * NB, m2, tid);
}

}  // namespace leaf32

/* ------------------------------------------------------------------------ */
/* FP64 leaf: two waves, column steps from an LDS-broadcast column           */
/* ------------------------------------------------------------------------ */
/*
 * The 64 x 64 FP64 leaf.  Wave 0's lane i holds row i of the tile
 * (a[k] = A(i, k)); wave 1's lane i holds column i of its inverse
 * (a[k] = X(k, i), started from e_i).  Step j: wave 0 publishes column j
 * (every lane's a[j]) in LDS, one barrier, and both waves read it back as
 * broadcasts c[k] = A(k, j) and apply the same column step
 *     a[j] /= L(j, j),   a[k] -= (a[j] / L(j, j)) c[k]   (k > j),
 * which for a row is  L(i, j) = A(i, j) / L(j, j),  A(i, k) -= L(i, j) L(k, j)
 * (right-looking Cholesky) and for an inverse column the column-oriented
 * forward substitution of L x = e_i.  The column buffer is double-buffered,
 * so one barrier per step orders both the publication and the reuse.  A
 * ragged leaf (m < 64) is padded with the identity.  1 / L(j, j) is the
 * hardware reciprocal square root of the pivot refined by Newton steps.
 */
namespace leaf64 {

constexpr int NB = 64, P = NB + 1;

__global__ __launch_bounds__(128) void potrf_leaf64_kernel(double *base, RfpView v, int m, int *info,
                                                           int info_base, double *inv)
{
    __shared__ double s[NB * P];
    __shared__ __attribute__((aligned(16))) double W[2][NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (*info != INT_MAX) return;          /* an earlier leaf already failed */

    /* the lower triangle into s, zero above, identity beyond m; loads run
     * along the view's contiguous direction.  Every load is issued before
     * the first LDS store: an element outside the lower triangle of the
     * m x m tile is loaded from (0, 0) instead (always inside the view) and
     * not used, so no load sits in a branch of its own. */
    double x[NB / 2];
#pragma unroll
    for (int e = 0; e < NB / 2; ++e) {
        const int o = 2 * e + wave;        /* the other index, two per step */
        const int r = v.rs == 1 ? lane : o, c = v.rs == 1 ? o : lane;
        const bool in = r < m && c < m && r >= c;
        x[e] = base[v.off + (in ? (int64_t)r * v.rs + (int64_t)c * v.cs : 0)];
    }
#pragma unroll
    for (int e = 0; e < NB / 2; ++e) {
        const int o = 2 * e + wave;
        const int r = v.rs == 1 ? lane : o, c = v.rs == 1 ? o : lane;
        double val = r == c ? 1.0 : 0.0;
        if (r < m && c < m) val = r >= c ? x[e] : 0.0;
        s[r * P + c] = val;
    }
    __syncthreads();
    double a[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) a[k] = wave == 0 ? s[lane * P + k] : (k == lane ? 1.0 : 0.0);

    /* The pivot test is a select (uniform: both waves read the same pivot),
     * not an exit, so the column loop has no branches; after a failing pivot
     * the arithmetic runs on, and only the columns before it are kept. */
    int fail = -1;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (wave == 0) W[j & 1][lane] = a[j];
        __syncthreads();
        const double piv = W[j & 1][j];
        fail = (fail < 0 && !(piv > 0.0)) ? j : fail;
        /* 1 / sqrt(pivot): the hardware estimate and two Newton steps
         * r += r (1 - piv r^2) / 2 (each doubles the correct bits); L(j, j) =
         * piv r.  A few dependent FMAs instead of an IEEE sqrt and division on
         * the column chain, to within an ulp or two of them. */
        double rinv = __builtin_amdgcn_rsq(piv);
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const double e = fma(-piv * rinv, rinv, 1.0);
            rinv = fma(0.5 * rinv, e, rinv);
        }
        const double ljj = piv * rinv;
        const double aj = a[j];
        a[j] = (wave == 0 && lane == j) ? ljj : aj * rinv;
        const double t = a[j] * rinv;
#pragma unroll
        for (int k = j + 1; k < NB; ++k) a[k] = fma(-t, W[j & 1][k], a[k]);
    }

    /* L back to the view through s (coalesced), the inverse straight to its
     * row-major slot (lane i writes column i); after a failure the columns
     * factored before it are written back and info set */
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < NB; ++k) s[lane * P + k] = a[k];
    }
    __syncthreads();
    const int cend = fail >= 0 ? fail : m;
    {
        double y[NB / 2];
#pragma unroll
        for (int e = 0; e < NB / 2; ++e) {
            const int o = 2 * e + wave;
            const int r = v.rs == 1 ? lane : o, c = v.rs == 1 ? o : lane;
            y[e] = s[r * P + c];
        }
#pragma unroll
        for (int e = 0; e < NB / 2; ++e) {
            const int o = 2 * e + wave;
            const int r = v.rs == 1 ? lane : o, c = v.rs == 1 ? o : lane;
            if (r < m && c < cend && r >= c) base[v.off + (int64_t)r * v.rs + (int64_t)c * v.cs] = y[e];
        }
    }
    if (fail >= 0) {
        if (tid == 0) atomicMin(info, info_base + fail + 1);
        return;
    }
    if (inv && wave == 1) {
#pragma unroll
        for (int k = 0; k < NB; ++k) inv[k * NB + lane] = (k < m && lane < m) ? a[k] : 0.0;
    }
}

}  // namespace leaf64

template <typename T>
static hipError_t potrf_tile(T *base, RfpView v, int m, int *info, int info_base, T *inv, hipStream_t s)
{
    constexpr int NB = LeafSize<T>::value;
    if (m <= 0) return hipSuccess;
    if (m > NB) return hipErrorInvalidValue;
    if (sizeof(T) == 4)
        hipLaunchKernelGGL(leaf32::potrf_leaf_kernel, dim3(1), dim3(256), 0, s, (float *)base, v, m, info, info_base,
                           (float *)inv);
    else
        hipLaunchKernelGGL(leaf64::potrf_leaf64_kernel, dim3(1), dim3(128), 0, s, (double *)base, v, m, info,
                           info_base, (double *)inv);
    return hipGetLastError();
}

template <typename T>
static hipError_t trtri_tiles(const T *base, RfpView v1, int64_t m1, T *inv1, RfpView v2, int64_t m2, T *inv2,
                              hipStream_t s)
{
    constexpr int NB = LeafSize<T>::value;
    const int64_t nt1 = m1 > 0 ? (m1 + NB - 1) / NB : 0, nt2 = m2 > 0 ? (m2 + NB - 1) / NB : 0;
    if (nt1 + nt2 == 0) return hipSuccess;
    hipLaunchKernelGGL((trtri_tiles_kernel<T, NB>), dim3((unsigned)(nt1 + nt2)), dim3(256), 0, s, base, v1, m1, inv1,
                       nt1, v2, m2, inv2);
    return hipGetLastError();
}

hipError_t launch_potrf_tile_f32(float *base, RfpView v, int m, int *info, int info_base, float *inv, hipStream_t s)
{ return potrf_tile<float>(base, v, m, info, info_base, inv, s); }

int potrf_pair_max() { return 2 * LeafSize<float>::value; }

hipError_t launch_potrf_pair_f32(float *base, RfpView v, int m, int *info, int info_base, float *inv,
                                 unsigned short *sh_hi, unsigned short *sh_lo, RfpView sv, bool bf16, hipStream_t s)
{
    constexpr int NB = LeafSize<float>::value;
    if (m <= NB || m > 2 * NB || !inv) return hipErrorInvalidValue;
    hipLaunchKernelGGL(leaf32::potrf_pair_kernel, dim3(1), dim3(256), 0, s, base, v, m, info, info_base, inv, sh_hi,
                       sh_lo, sv, bf16 ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_potrf_tile_f64(double *base, RfpView v, int m, int *info, int info_base, double *inv, hipStream_t s)
{ return potrf_tile<double>(base, v, m, info, info_base, inv, s); }
hipError_t launch_trtri_tiles_f32(const float *base, RfpView v, int64_t m, float *inv, hipStream_t s)
{ return trtri_tiles<float>(base, v, m, inv, v, 0, inv, s); }
hipError_t launch_trtri_tiles_f64(const double *base, RfpView v, int64_t m, double *inv, hipStream_t s)
{ return trtri_tiles<double>(base, v, m, inv, v, 0, inv, s); }
hipError_t launch_trtri_tiles2_f32(const float *base, RfpView v1, int64_t m1, float *inv1, RfpView v2, int64_t m2,
                                   float *inv2, hipStream_t s)
{ return trtri_tiles<float>(base, v1, m1, inv1, v2, m2, inv2, s); }
hipError_t launch_trtri_tiles2_f64(const double *base, RfpView v1, int64_t m1, double *inv1, RfpView v2, int64_t m2,
                                   double *inv2, hipStream_t s)
{ return trtri_tiles<double>(base, v1, m1, inv1, v2, m2, inv2, s); }
