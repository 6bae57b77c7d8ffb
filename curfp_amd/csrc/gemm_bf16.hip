/*
 * gemm_bf16.hip — BF16 and split-BF16 MFMA variants of the view GEMM
 * (gemm.hip):
 *
 *   C(i, j) = alpha * sum_k A(i, k) * B(j, k) + beta * C(i, j)
 *
 * used for ssfrk and the spftrf SYRK / GEMM updates when the handle's math
 * mode is BF16 or BF16X3, and for ssfrk with a BF16 input matrix.
 *
 *   BF16   : operands rounded to BF16 (round-to-nearest-even) while they are
 *            staged into LDS; v_mfma_f32_32x32x16_bf16, FP32 accumulation.
 *   BF16X3 : x = hi + lo with hi = bf16(x), lo = bf16(x - hi); the product is
 *            hi*hi + hi*lo + lo*hi (three MFMAs), which keeps ~16 mantissa
 *            bits of every product — close to FP32 products at BF16 rate / 3.
 *
 * Structure: 256 threads (4 wave64s, 2 x 2), 128 x 128 output tile, BK = 32;
 * every thread stages one 4 x 4 block of each operand per k-slice (loaded
 * along the view's contiguous direction, transposed in registers) so LDS
 * holds rows with k contiguous, and each lane feeds the MFMA from one
 * 16-byte LDS read.  LDS is double-buffered with the next slice's global
 * loads in flight during the MFMAs.
 */
#include "rfp_internal.h"
#include "mfma_epilogue.h"

typedef float  f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned short u16x4u __attribute__((ext_vector_type(4), aligned(2)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

namespace {

struct BKParams {
    const void *a; int64_t a_off, a_rs, a_cs;
    const void *b; int64_t b_off, b_rs, b_cs;
    float *c;      int64_t c_off, c_rs, c_cs;
    int64_t M, N, K;
    float alpha, beta;
    int64_t tiles_m, tiles_n, ntiles;
    int lower;
};

__device__ __forceinline__ int64_t xcd_remap(int64_t pid, int64_t total)
{
    const int64_t X = 8;
    const int64_t q = total / X, r = total % X;
    const int64_t x = pid % X, w = pid / X;
    return x * q + (x < r ? x : r) + w;
}

__device__ __forceinline__ void tile_coords(const BKParams &p, int64_t t, int64_t &tm, int64_t &tn)
{
    if (p.lower) {
        tri_tile_coords(t, p.tiles_m, tm, tn);
    } else {
        const int64_t GROUP = 8;
        const int64_t in_group = GROUP * p.tiles_n;
        const int64_t gid = t / in_group;
        const int64_t first = gid * GROUP;
        const int64_t gsz = (p.tiles_m - first) < GROUP ? (p.tiles_m - first) : GROUP;
        const int64_t loc = t % in_group;
        tm = first + loc % gsz;
        tn = loc / gsz;
    }
}

/* storage element -> float */
__device__ __forceinline__ float to_f(float x) { return x; }
__device__ __forceinline__ float to_f(unsigned short x) { return __uint_as_float((unsigned)x << 16); }

__device__ __forceinline__ unsigned short bf16_bits(float x)
{
    __bf16 b = (__bf16)x;   /* round to nearest even */
    return __builtin_bit_cast(unsigned short, b);
}

template <typename S> struct SVec;
template <> struct SVec<float> {
    __device__ static void load4(const float *p, float *o) { f32x4u v = *(const f32x4u *)p; o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3]; }
};
template <> struct SVec<unsigned short> {
    __device__ static void load4(const unsigned short *p, float *o)
    {
        u16x4u v = *(const u16x4u *)p;
        o[0] = to_f((unsigned short)v[0]); o[1] = to_f((unsigned short)v[1]);
        o[2] = to_f((unsigned short)v[2]); o[3] = to_f((unsigned short)v[3]);
    }
};

/* One 4 x 4 block (rows r..r+3, k kq..kq+3) per thread: ROWS * BK == 16 * 256. */
template <typename S, int ROWS, int BK, bool KC>
struct Stage4x4 {
    float v[4][4];  /* v[row][k] */
    int r, kq;

    __device__ __forceinline__ void coords(int tid)
    {
        if (KC) { r = (tid / (BK / 4)) * 4; kq = (tid % (BK / 4)) * 4; }
        else    { r = (tid % (ROWS / 4)) * 4; kq = (tid / (ROWS / 4)) * 4; }
    }

    __device__ __forceinline__ void load(const S *base, int64_t off, int64_t rs, int64_t cs,
                                         int64_t r0, int64_t k0, int64_t R, int64_t K)
    {
        const int64_t gr = r0 + r, gk = k0 + kq;
        const bool full = gr + 4 <= R && gk + 4 <= K;
        if (KC) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const S *p = base + off + (gr + c) * rs + gk;
                if (full) {
                    SVec<S>::load4(p, v[c]);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[c][e] = (gr + c < R && gk + e < K) ? to_f(p[e]) : 0.f;
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const S *p = base + off + gr + (gk + e) * cs;
                float t[4];
                if (full) {
                    SVec<S>::load4(p, t);
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c) t[c] = (gr + c < R && gk + e < K) ? to_f(p[c]) : 0.f;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) v[c][e] = t[c];
            }
        }
    }

    /* rows with k contiguous, pitch P bf16 elements */
    template <int P, bool X3>
    __device__ __forceinline__ void store(unsigned short *hi, unsigned short *lo) const
    {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            u16x4 h, l;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const unsigned short hb = bf16_bits(v[c][e]);
                h[e] = hb;
                if (X3) l[e] = bf16_bits(v[c][e] - to_f(hb));
            }
            *(u16x4 *)(hi + (r + c) * P + kq) = h;
            if (X3) *(u16x4 *)(lo + (r + c) * P + kq) = l;
        }
    }
};

template <typename S, bool X3, bool AKC, bool BKC>
__global__ __launch_bounds__(256) void view_gemm_bf16_kernel(BKParams p)
{
    constexpr int BM = 128, BN = 128, BK = 32, P = BK + 8;
    constexpr int NB = X3 ? 2 : 1;
    __shared__ __attribute__((aligned(16))) unsigned short sA[2][NB][BM * P];
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][NB][BN * P];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    int64_t tm, tn;
    tile_coords(p, xcd_remap(blockIdx.x, p.ntiles), tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const S *A = (const S *)p.a;
    const S *B = (const S *)p.b;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    Stage4x4<S, BM, BK, AKC> la;
    Stage4x4<S, BN, BK, BKC> lb;
    la.coords(tid);
    lb.coords(tid);
    const int64_t nk = (p.K + BK - 1) / BK;
    if (nk > 0) {
        la.load(A, p.a_off, p.a_rs, p.a_cs, m0, 0, p.M, p.K);
        lb.load(B, p.b_off, p.b_rs, p.b_cs, n0, 0, p.N, p.K);
        la.template store<P, X3>(sA[0][0], sA[0][NB - 1]);
        lb.template store<P, X3>(sB[0][0], sB[0][NB - 1]);
    }
    __syncthreads();

    const int lrow = lane & 31, lk = (lane >> 5) * 8;
    for (int64_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            la.load(A, p.a_off, p.a_rs, p.a_cs, m0, (kt + 1) * BK, p.M, p.K);
            lb.load(B, p.b_off, p.b_rs, p.b_cs, n0, (kt + 1) * BK, p.N, p.K);
        }
#pragma unroll
        for (int k = 0; k < BK; k += 16) {
            bf16x8 ah[2], bh[2], al[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int off = (wm * 64 + i * 32 + lrow) * P + k + lk;
                ah[i] = *(const bf16x8 *)(sA[cur][0] + off);
                if (X3) al[i] = *(const bf16x8 *)(sA[cur][NB - 1] + off);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int off = (wn * 64 + j * 32 + lrow) * P + k + lk;
                bh[j] = *(const bf16x8 *)(sB[cur][0] + off);
                if (X3) bl[j] = *(const bf16x8 *)(sB[cur][NB - 1] + off);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (X3) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    }
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                }
        }
        if (more) {
            la.template store<P, X3>(sA[cur ^ 1][0], sA[cur ^ 1][NB - 1]);
            lb.template store<P, X3>(sB[cur ^ 1][0], sB[cur ^ 1][NB - 1]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            store_block32(acc[i][j], p.c, p.c_off, p.c_rs, p.c_cs, m0 + wm * 64 + i * 32, n0 + wn * 64 + j * 32,
                          p.M, p.N, p.lower != 0, p.alpha, p.beta, lane);
}

template <typename S, bool X3>
hipError_t launch_typed(const BKParams &p, bool akc, bool bkc, hipStream_t s)
{
    dim3 grid((unsigned)p.ntiles), block(256);
    if (akc && bkc)  hipLaunchKernelGGL((view_gemm_bf16_kernel<S, X3, true, true>), grid, block, 0, s, p);
    else if (akc)    hipLaunchKernelGGL((view_gemm_bf16_kernel<S, X3, true, false>), grid, block, 0, s, p);
    else if (bkc)    hipLaunchKernelGGL((view_gemm_bf16_kernel<S, X3, false, true>), grid, block, 0, s, p);
    else             hipLaunchKernelGGL((view_gemm_bf16_kernel<S, X3, false, false>), grid, block, 0, s, p);
    return hipGetLastError();
}

/*
 * BF16-operand kernel for the factorisation's updates.  A and B are BF16
 * row blocks with k contiguous (the factor's BF16 shadow, see
 * rfp_drivers.hip): A(i, k) = a[a_off + i*a_ld + k], rows 16-byte aligned.
 * Tile BT x BT (BT = 256: 4 waves of 128 x 128, 4 x 4 MFMA blocks each, 256
 * accumulator registers per lane; BT = 128 for split-BF16, whose hi + lo
 * operands double the LDS; BT = 64 for small problems), BK = 32 or 64,
 * double-buffered LDS, one 16-byte global
 * load and one 16-byte LDS store per vector, one 16-byte LDS read per MFMA
 * operand.
 */
struct RowParams {
    const unsigned short *a_hi, *a_lo; int64_t a_off, a_ld;
    const unsigned short *b_hi, *b_lo; int64_t b_off, b_ld;
    float *c; int64_t c_off, c_rs, c_cs;
    int64_t M, N, K;
    float alpha, beta;
    int64_t tiles_m, tiles_n, ntiles;
    int lower;
    int batch;                      /* diagonal blocks along blockIdx.y (GemmArgs::batch) */
    int64_t a_bs, b_bs, c_bs, m_last;
};

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <int BT, int BK>
struct RowStage {
    static constexpr int NV = BT * BK / 8 / 256;   /* 16-byte vectors per thread */
    u16x8 v[NV];

    __device__ __forceinline__ void load(const unsigned short *base, int64_t off, int64_t ld, int64_t r0, int64_t k0,
                                         int64_t R, int64_t K, int tid)
    {
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int id = tid + 256 * e;
            const int row = id / (BK / 8), kq = (id % (BK / 8)) * 8;
            const int64_t gr = r0 + row, gk = k0 + kq;
            const unsigned short *p = base + off + gr * ld + gk;
            if (gr < R && gk + 8 <= K) {
                v[e] = *(const u16x8 *)p;
            } else {
                u16x8 t;
#pragma unroll
                for (int c = 0; c < 8; ++c) t[c] = (gr < R && gk + c < K) ? p[c] : (unsigned short)0;
                v[e] = t;
            }
        }
    }

    template <int P>
    __device__ __forceinline__ void store(unsigned short *s, int tid) const
    {
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int id = tid + 256 * e;
            const int row = id / (BK / 8), kq = (id % (BK / 8)) * 8;
            *(u16x8 *)(s + row * P + kq) = v[e];
        }
    }
};

/* SWAP: the MFMA operands exchanged (acc holds C's blocks transposed) for
 * row-view outputs, whose epilogue then reads and writes 16-byte runs */
template <int BT, bool X3, int BK, bool SWAP>
__global__ __launch_bounds__(256) void gemm_bf16_rows_kernel(RowParams p)
{
    constexpr int P = BK + 8;
    constexpr int WT = BT / 2, FM = WT / 32;
    constexpr int NS = X3 ? 2 : 1;
    __shared__ __attribute__((aligned(16))) unsigned short sA[2][NS][BT * P];
    __shared__ __attribute__((aligned(16))) unsigned short sB[2][NS][BT * P];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    if (p.batch > 1) {
        /* diagonal block bz of a batched lower-triangular launch */
        const int64_t bz = blockIdx.y;
        p.a_off += bz * p.a_bs;
        p.b_off += bz * p.b_bs;
        p.c_off += bz * p.c_bs;
        if (bz == p.batch - 1) p.M = p.N = p.m_last;
    }
    int64_t tm, tn;
    {
        BKParams q;
        q.tiles_m = p.tiles_m; q.tiles_n = p.tiles_n; q.lower = p.lower;
        tile_coords(q, xcd_remap(blockIdx.x, p.ntiles), tm, tn);
    }
    const int64_t m0 = tm * BT, n0 = tn * BT;
    if (m0 >= p.M || n0 >= p.N) return;   /* past a smaller last block (whole workgroup) */

    f32x16 acc[FM][FM];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FM; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    RowStage<BT, BK> ah, bh, al, bl;
    const int64_t nk = (p.K + BK - 1) / BK;
    if (nk > 0) {
        ah.load(p.a_hi, p.a_off, p.a_ld, m0, 0, p.M, p.K, tid);
        bh.load(p.b_hi, p.b_off, p.b_ld, n0, 0, p.N, p.K, tid);
        if (X3) {
            al.load(p.a_lo, p.a_off, p.a_ld, m0, 0, p.M, p.K, tid);
            bl.load(p.b_lo, p.b_off, p.b_ld, n0, 0, p.N, p.K, tid);
        }
        ah.template store<P>(sA[0][0], tid);
        bh.template store<P>(sB[0][0], tid);
        if (X3) {
            al.template store<P>(sA[0][NS - 1], tid);
            bl.template store<P>(sB[0][NS - 1], tid);
        }
    }
    __syncthreads();

    const int lrow = lane & 31, lk = (lane >> 5) * 8;
    for (int64_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            const int64_t k0 = (kt + 1) * BK;
            /* the next slice's loads stay in flight during this slice's MFMAs */
            ah.load(p.a_hi, p.a_off, p.a_ld, m0, k0, p.M, p.K, tid);
            bh.load(p.b_hi, p.b_off, p.b_ld, n0, k0, p.N, p.K, tid);
            if (X3) {
                al.load(p.a_lo, p.a_off, p.a_ld, m0, k0, p.M, p.K, tid);
                bl.load(p.b_lo, p.b_off, p.b_ld, n0, k0, p.N, p.K, tid);
            }
        }
#pragma unroll 1
        for (int k = 0; k < BK; k += 16) {
            bf16x8 fa[FM], fb[FM], fal[FM], fbl[FM];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int off = (wm * WT + i * 32 + lrow) * P + k + lk;
                fa[i] = *(const bf16x8 *)(sA[cur][0] + off);
                if (X3) fal[i] = *(const bf16x8 *)(sA[cur][NS - 1] + off);
            }
#pragma unroll
            for (int j = 0; j < FM; ++j) {
                const int off = (wn * WT + j * 32 + lrow) * P + k + lk;
                fb[j] = *(const bf16x8 *)(sB[cur][0] + off);
                if (X3) fbl[j] = *(const bf16x8 *)(sB[cur][NS - 1] + off);
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FM; ++j) {
                    if (SWAP) {
                        if (X3) {
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[j], fal[i], acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fbl[j], fa[i], acc[i][j], 0, 0, 0);
                        }
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[j], fa[i], acc[i][j], 0, 0, 0);
                    } else {
                        if (X3) {
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fal[i], fb[j], acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fbl[j], acc[i][j], 0, 0, 0);
                        }
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
                    }
                }
        }
        if (more) {
            ah.template store<P>(sA[cur ^ 1][0], tid);
            bh.template store<P>(sB[cur ^ 1][0], tid);
            if (X3) {
                al.template store<P>(sA[cur ^ 1][NS - 1], tid);
                bl.template store<P>(sB[cur ^ 1][NS - 1], tid);
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FM; ++j) {
            if (SWAP)
                store_block32_t(acc[i][j], p.c, p.c_off, p.c_rs, p.c_cs, m0 + wm * WT + i * 32, n0 + wn * WT + j * 32,
                                p.M, p.N, p.lower != 0, p.alpha, p.beta, lane);
            else
                store_block32(acc[i][j], p.c, p.c_off, p.c_rs, p.c_cs, m0 + wm * WT + i * 32, n0 + wn * WT + j * 32,
                              p.M, p.N, p.lower != 0, p.alpha, p.beta, lane);
        }
}

template <int BT, bool X3, int BK>
hipError_t launch_rows(const GemmArgs &g, hipStream_t s)
{
    RowParams p;
    p.a_hi = (const unsigned short *)g.a; p.a_lo = (const unsigned short *)g.a_lo;
    p.a_off = g.va.off; p.a_ld = g.va.rs;
    p.b_hi = (const unsigned short *)g.b; p.b_lo = (const unsigned short *)g.b_lo;
    p.b_off = g.vb.off; p.b_ld = g.vb.rs;
    p.c = (float *)g.c; p.c_off = g.vc.off; p.c_rs = g.vc.rs; p.c_cs = g.vc.cs;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.alpha = (float)g.alpha; p.beta = (float)g.beta;
    p.lower = g.lower ? 1 : 0;
    p.batch = g.batch > 1 ? g.batch : 1;
    p.a_bs = g.a_bs; p.b_bs = g.b_bs; p.c_bs = g.c_bs; p.m_last = g.m_last;
    if (p.batch > 1 && (!g.lower || g.m_last > g.M || g.m_last < 1)) return hipErrorInvalidValue;
    p.tiles_m = (g.M + BT - 1) / BT;
    p.tiles_n = (g.N + BT - 1) / BT;
    p.ntiles = g.lower ? p.tiles_m * (p.tiles_m + 1) / 2 : p.tiles_m * p.tiles_n;
    if (p.ntiles <= 0) return hipSuccess;
    if (p.ntiles > 0x7fffffffLL) return hipErrorInvalidValue;
    /* row view read and written (beta != 0): transposed accumulators, 16-byte runs along rows */
    const bool swap = g.vc.cs == 1 && g.vc.rs != 1 && g.beta != 0.0;
    const dim3 grid((unsigned)p.ntiles, (unsigned)p.batch);
    if (swap) hipLaunchKernelGGL((gemm_bf16_rows_kernel<BT, X3, BK, true>), grid, dim3(256), 0, s, p);
    else      hipLaunchKernelGGL((gemm_bf16_rows_kernel<BT, X3, BK, false>), grid, dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace

namespace {

/* out[i * ldo + p] = bf16(A(i, p)) for p < K, 0 for K <= p < Kp (and the
 * split-BF16 low part bf16(A - hi) into lo when lo is given); A(i, p) at
 * off + i * rs + p * cs.  Consecutive threads run along A's contiguous
 * direction. */
__global__ __launch_bounds__(256) void to_bf16_rows_kernel(const float *A, RfpView av, int64_t rows, int64_t K,
                                                           int64_t Kp, int64_t ldo, unsigned short *hi,
                                                           unsigned short *lo)
{
    /* a thread converts 4 consecutive k of a row: x over the k groups, y
     * over the rows (strided; a short Kp puts several rows in a workgroup);
     * 16-byte loads where k is contiguous and the group lies inside K,
     * 8-byte stores where the output rows allow */
    typedef float f32x4a __attribute__((ext_vector_type(4), aligned(4)));
    typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
    const int64_t p4 = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (p4 >= Kp) return;
    const bool vec_in = av.cs == 1 && p4 + 4 <= K;
    const bool vec_out = ldo % 4 == 0 && p4 + 4 <= Kp && ((uintptr_t)hi & 7) == 0 && ((uintptr_t)lo & 7) == 0;
    for (int64_t i = (int64_t)blockIdx.y * blockDim.y + threadIdx.y; i < rows; i += (int64_t)gridDim.y * blockDim.y) {
        const float *src = A + av.off + i * av.rs + p4 * av.cs;
        float x[4];
        if (vec_in) {
            const f32x4a t = *(const f32x4a *)src;
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = t[u];
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = p4 + u < K ? src[u * av.cs] : 0.f;
        }
        u16x4 h, l;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            h[u] = bf16_bits(x[u]);
            l[u] = bf16_bits(x[u] - to_f((unsigned short)h[u]));
        }
        unsigned short *ho = hi + i * ldo + p4, *lw = lo ? lo + i * ldo + p4 : nullptr;
        if (vec_out) {
            *(u16x4 *)ho = h;
            if (lw) *(u16x4 *)lw = l;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (p4 + u < Kp) {
                    ho[u] = h[u];
                    if (lw) lw[u] = l[u];
                }
        }
    }
}

/* The same for an operand whose rows are contiguous (cs != 1, rs == 1): a
 * 64 x 64 tile goes through LDS, read along rows and written along k. */
__global__ __launch_bounds__(256) void to_bf16_rows_t_kernel(const float *A, RfpView av, int64_t rows, int64_t K,
                                                             int64_t Kp, int64_t ldo, unsigned short *hi,
                                                             unsigned short *lo)
{
    __shared__ float tile[64][65];
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * 64, i0 = (int64_t)blockIdx.y * 64;
#pragma unroll 4
    for (int e = 0; e < 16; ++e) {
        const int ii = tid % 64, pp = tid / 64 + 4 * e;
        const int64_t i = i0 + ii, p = p0 + pp;
        tile[pp][ii] = (i < rows && p < K) ? A[av.off + i * av.rs + p * av.cs] : 0.f;
    }
    __syncthreads();
#pragma unroll 4
    for (int e = 0; e < 16; ++e) {
        const int pp = tid % 64, ii = tid / 64 + 4 * e;
        const int64_t i = i0 + ii, p = p0 + pp;
        if (i < rows && p < Kp) {
            const float x = tile[pp][ii];
            const unsigned short h = bf16_bits(x);
            hi[i * ldo + p] = h;
            if (lo) lo[i * ldo + p] = bf16_bits(x - to_f(h));
        }
    }
}

}  // namespace

hipError_t launch_to_bf16_rows(const float *A, RfpView av, int64_t rows, int64_t K, int64_t Kp, int64_t ldo,
                               unsigned short *hi, unsigned short *lo, hipStream_t s)
{
    if (rows <= 0 || Kp <= 0) return hipSuccess;
    if (K > Kp || Kp > ldo) return hipErrorInvalidValue;
    if (av.cs != 1 && av.rs == 1 && (rows + 63) / 64 <= 65535) {
        const dim3 grid((unsigned)((Kp + 63) / 64), (unsigned)((rows + 63) / 64));
        hipLaunchKernelGGL(to_bf16_rows_t_kernel, grid, dim3(256), 0, s, A, av, rows, K, Kp, ldo, hi, lo);
        return hipGetLastError();
    }
    unsigned bx = 1;
    while (bx < 256 && 4 * (int64_t)bx < Kp) bx *= 2;
    const dim3 block(bx, 256 / bx);
    const dim3 grid((unsigned)((Kp + 4 * bx - 1) / (4 * bx)),
                    (unsigned)std::min<int64_t>((rows + block.y - 1) / block.y, 16384));
    hipLaunchKernelGGL(to_bf16_rows_kernel, grid, block, 0, s, A, av, rows, K, Kp, ldo, hi, lo);
    return hipGetLastError();
}

hipError_t launch_gemm_bf16(const GemmArgs &g, RfpMath math, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    if (g.inplace || g.splits > 1) return hipErrorInvalidValue;
    if (g.bf16_rows) {
        /* BF16 row blocks (k contiguous, 16-byte aligned rows) */
        if (g.va.cs != 1 || g.vb.cs != 1 || (g.va.rs % 8) || (g.vb.rs % 8) || (g.va.off % 8) || (g.vb.off % 8))
            return hipErrorInvalidValue;
        if (g.lower && g.M != g.N) return hipErrorInvalidValue;
        const bool x3 = math == RFP_MATH_BF16X3;
        if (x3 && (!g.a_lo || !g.b_lo)) return hipErrorInvalidValue;
        /* 256 x 256 tiles for the large updates (M, N >= 4096); 64 x 64
         * tiles (BK = 64) for problems of fewer than 256 tiles of 128 x 128,
         * where a workgroup's own latency sets the time (four times the
         * workgroups, a quarter of the work and half the k-slices each).
         * Both thresholds measured on the spftrf chain at n = 65536. */
        const int64_t tm = (g.M + 127) / 128, tn = (g.N + 127) / 128;
        const bool small = (g.lower ? tm * (tm + 1) / 2 : tm * tn) * (g.batch > 1 ? g.batch : 1) < 256;
        const bool big = g.M >= 4096 && g.N >= 4096;
        if (x3) return small ? launch_rows<64, true, 64>(g, s) : launch_rows<128, true, 32>(g, s);
        if (small) return launch_rows<64, false, 64>(g, s);
        const hipError_t e = launch_gemm_bf16_dma(g, big ? 256 : 128, s);
        if (e != hipErrorNotSupported) return e;
        return big ? launch_rows<256, false, 64>(g, s) : launch_rows<128, false, 32>(g, s);
    }
    if (g.batch > 1) return hipErrorInvalidValue;   /* batched launches take the row-block kernels */
    BKParams p;
    p.a = g.a; p.a_off = g.va.off; p.a_rs = g.va.rs; p.a_cs = g.va.cs;
    p.b = g.b; p.b_off = g.vb.off; p.b_rs = g.vb.rs; p.b_cs = g.vb.cs;
    p.c = (float *)g.c; p.c_off = g.vc.off; p.c_rs = g.vc.rs; p.c_cs = g.vc.cs;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.alpha = (float)g.alpha; p.beta = (float)g.beta;
    p.lower = g.lower ? 1 : 0;
    p.tiles_m = (g.M + 127) / 128;
    p.tiles_n = (g.N + 127) / 128;
    if (g.lower) {
        if (g.M != g.N) return hipErrorInvalidValue;
        p.ntiles = p.tiles_m * (p.tiles_m + 1) / 2;
    } else {
        p.ntiles = p.tiles_m * p.tiles_n;
    }
    if (p.ntiles <= 0) return hipSuccess;
    if (p.ntiles > 0x7fffffffLL) return hipErrorInvalidValue;
    const bool akc = g.va.cs == 1, bkc = g.vb.cs == 1;
    if (g.K > 0 && ((!akc && g.va.rs != 1) || (!bkc && g.vb.rs != 1))) return hipErrorInvalidValue;
    /* a BF16 input is exact in BF16: split-BF16 adds nothing there */
    const bool x3 = math == RFP_MATH_BF16X3 && !g.a_bf16;
    if (g.a_bf16) return launch_typed<unsigned short, false>(p, akc, bkc, s);
    if (x3) return launch_typed<float, true>(p, akc, bkc, s);
    return launch_typed<float, false>(p, akc, bkc, s);
}
