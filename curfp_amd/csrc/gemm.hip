/*
 * gemm.hip — MFMA "view GEMM" for gfx950: the one contraction kernel behind
 * ssfrk (SYRK + GEMM blocks), the spftrf updates (TRSM-by-inverse, SYRK,
 * GEMM) and the spftrs block substitutions.
 *
 *   C(i, j) = alpha * sum_k A(i, k) * B(j, k) + beta * C(i, j)
 *
 * A, B and C are strided views (element (i, j) at off + i*rs + j*cs) straight
 * into the packed RFP array or the caller's dense arrays, so no sub-block is
 * ever copied out of the rectangle.  The reference reaches the same blocks
 * through cuBLAS leading-dimension arguments (src/curfp_ssfrk.cpp:193-209,
 * src/curfp_spftrf.cpp:263-274, src/curfp_spftrs.cpp:258-299).
 *
 * Structure (one workgroup = 256 threads = 4 wave64s):
 *   - 128 x 128 (or 128 x 64) output tile, 2 x 2 (4 x 1) waves;
 *   - A and B k-slices (BK = 16) staged through LDS, double-buffered, the
 *     next slice's global loads in flight during the current slice's MFMAs;
 *     loads run along whichever dimension of the view is contiguous
 *     (16-byte vector loads), so both row and column views stream coalesced;
 *   - FP32: v_mfma_f32_32x32x2_f32; FP64: v_mfma_f64_16x16x4_f64;
 *   - lower-triangular outputs (SYRK) enumerate only the tiles on or below
 *     the diagonal; workgroup ids are remapped so that neighbouring tiles
 *     (which share A/B panels) land on the same XCD and its L2.
 */
#include "rfp_internal.h"
#include <algorithm>

typedef float  f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4  __attribute__((ext_vector_type(4)));

/* 16-byte vectors that only assume element alignment (RFP views have odd
 * leading dimensions such as n+1, so rows are not 16-byte aligned). */
typedef float  f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));

template <typename T> struct Vec;
template <> struct Vec<float>  { using type = f32x4u; static constexpr int N = 4; };
template <> struct Vec<double> { using type = f64x2u; static constexpr int N = 2; };

/* MFMA shape traits. */
template <typename T> struct Mfma;
template <> struct Mfma<float> {
    static constexpr int TM = 32, TN = 32, TK = 2, NACC = 16;
    using acc_t = f32x16;
    __device__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    /* output row of accumulator element r held by lane l (column = l % 32) */
    __device__ static int row(int l, int r) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
    __device__ static int col(int l) { return l & 31; }
    __device__ static int kk(int l) { return l >> 5; }
    __device__ static int mn(int l) { return l & 31; }
};
template <> struct Mfma<double> {
    static constexpr int TM = 16, TN = 16, TK = 4, NACC = 4;
    using acc_t = f64x4;
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    /* FP64 output: register r of lane l holds row (l / 16) + 4 r — unlike the
     * FP32 16x16 shapes, whose rows are 4 (l / 16) + r (checked on gfx950) */
    __device__ static int row(int l, int r) { return (l >> 4) + 4 * r; }
    __device__ static int col(int l) { return l & 15; }
    __device__ static int kk(int l) { return l >> 4; }
    __device__ static int mn(int l) { return l & 15; }
};

struct KParams {
    const void *a; int64_t a_off, a_rs, a_cs;
    const void *b; int64_t b_off, b_rs, b_cs;
    void *c;       int64_t c_off, c_rs, c_cs;
    int64_t M, N, K;
    double alpha, beta;
    int64_t tiles_m, tiles_n, ntiles;
    int lower;
    int64_t kchunk;     /* split-K slice length (== K without split-K) */
    int64_t c_zstride;  /* element stride between the partial C of two slices */
    unsigned short *sh_hi, *sh_lo;  /* optional BF16 copy of C (hi, lo parts) */
    int64_t sh_off, sh_rs, sh_cs;
    int batch;                      /* diagonal blocks along blockIdx.z (GemmArgs::batch) */
    int64_t a_bs, b_bs, c_bs, m_last;
};

__device__ __forceinline__ unsigned short bf16_rne(float x)
{
    __bf16 b = (__bf16)x;
    return __builtin_bit_cast(unsigned short, b);
}

/* Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
 * contiguous range of tile ids so neighbouring tiles share its L2. */
__device__ __forceinline__ int64_t xcd_remap(int64_t pid, int64_t total)
{
    const int64_t X = 8;
    const int64_t q = total / X, r = total % X;
    const int64_t x = pid % X, w = pid / X;
    return x * q + (x < r ? x : r) + w;
}

__device__ __forceinline__ void tile_coords(const KParams &p, int64_t t, int64_t &tm, int64_t &tn)
{
    if (p.lower) {
        tri_tile_coords(t, p.tiles_m, tm, tn);
    } else {
        /* grouped ordering: GROUP tile-rows walk the tile-columns together */
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

/*
 * Load a ROWS x BK slice (rows r0.., k k0..) of view (off, rs, cs) into
 * registers.  KC: the view is contiguous along k (cs == 1), otherwise along
 * rows (rs == 1).  Each thread holds NV vectors of Vec<S>::N elements in the
 * storage type S; the conversion to the compute type T happens in store(), so
 * that no load is waited for where it is issued (a conversion next to its
 * load would be).
 */
template <typename S, typename T, int ROWS, int BK, bool KC>
struct SliceLoader {
    static constexpr int VN = Vec<S>::N;
    static constexpr int NV = ROWS * BK / (256 * VN);
    static_assert(NV >= 1, "slice too small");
    S v[NV][VN];

    __device__ __forceinline__ void load(const S *base, int64_t off, int64_t rs, int64_t cs,
                                         int64_t r0, int64_t k0, int64_t R, int64_t K, int tid)
    {
        using V = typename Vec<S>::type;
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int id = tid + 256 * e;
            int rr, kq;
            if (KC) { rr = id / (BK / VN); kq = (id % (BK / VN)) * VN; }
            else    { rr = (id % (ROWS / VN)) * VN; kq = id / (ROWS / VN); }
            const int64_t gr = r0 + rr, gk = k0 + kq;
            if (KC) {
                const S *p = base + off + gr * rs + gk;
                if (gr < R && gk + VN <= K) {
                    V x = *(const V *)p;
#pragma unroll
                    for (int c = 0; c < VN; ++c) v[e][c] = x[c];
                } else {
#pragma unroll
                    for (int c = 0; c < VN; ++c) v[e][c] = (gr < R && gk + c < K) ? p[c] : S(0);
                }
            } else {
                const S *p = base + off + gr + gk * cs;
                if (gk < K && gr + VN <= R) {
                    V x = *(const V *)p;
#pragma unroll
                    for (int c = 0; c < VN; ++c) v[e][c] = x[c];
                } else {
#pragma unroll
                    for (int c = 0; c < VN; ++c) v[e][c] = (gk < K && gr + c < R) ? p[c] : S(0);
                }
            }
        }
    }

    /* store into LDS laid out s[k][row] (row contiguous, row pitch LDS) */
    template <int LDS>
    __device__ __forceinline__ void store(T *s, int tid) const
    {
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int id = tid + 256 * e;
            int rr, kq;
            if (KC) { rr = id / (BK / VN); kq = (id % (BK / VN)) * VN; }
            else    { rr = (id % (ROWS / VN)) * VN; kq = id / (ROWS / VN); }
#pragma unroll
            for (int c = 0; c < VN; ++c) {
                if (KC) s[(kq + c) * LDS + rr] = (T)v[e][c];
                else    s[kq * LDS + rr + c] = (T)v[e][c];
            }
        }
    }
};

template <typename T, typename AS, typename BS, int BM, int BN, int WAVES_M, int BK, bool AKC, bool BKC>
__global__ __launch_bounds__(256) void view_gemm_kernel(KParams p)
{
    using MF = Mfma<T>;
    constexpr int WAVES_N = 4 / WAVES_M;
    constexpr int WTM = BM / WAVES_M, WTN = BN / WAVES_N;
    constexpr int FM = WTM / MF::TM, FN = WTN / MF::TN;
    static_assert(FM >= 1 && FN >= 1, "wave tile too small");
    /* pitch chosen so the k-contiguous LDS stores are bank-conflict free */
    constexpr int LDA = BM + 4, LDB = BN + 4;
    __shared__ T sA[2][BK * LDA];
    __shared__ T sB[2][BK * LDB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    if (p.batch > 1) {
        /* diagonal block bz of a batched lower-triangular launch */
        const int64_t bz = blockIdx.z;
        p.a_off += bz * p.a_bs;
        p.b_off += bz * p.b_bs;
        p.c_off += bz * p.c_bs;
        if (bz == p.batch - 1) p.M = p.N = p.m_last;
    }
    int64_t tm, tn;
    tile_coords(p, xcd_remap(blockIdx.x, p.ntiles), tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;
    if (m0 >= p.M || n0 >= p.N) return;   /* past a smaller last block (whole workgroup) */

    /* split-K: slice z covers k in [k_lo, k_hi) and writes its own partial C */
    const int64_t z = blockIdx.y;
    const int64_t k_lo = z * p.kchunk;
    const int64_t k_hi = (k_lo + p.kchunk) < p.K ? (k_lo + p.kchunk) : p.K;
    const int64_t klen = k_hi > k_lo ? k_hi - k_lo : 0;

    const AS *A = (const AS *)p.a + k_lo * (AKC ? 1 : p.a_cs);
    const BS *B = (const BS *)p.b + k_lo * (BKC ? 1 : p.b_cs);
    T *C = (T *)p.c + z * p.c_zstride;

    typename MF::acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < MF::NACC; ++r) acc[i][j][r] = 0;

    const int64_t nk = (klen + BK - 1) / BK;
    SliceLoader<AS, T, BM, BK, AKC> la;
    SliceLoader<BS, T, BN, BK, BKC> lb;
    if (nk > 0) {
        la.load(A, p.a_off, p.a_rs, p.a_cs, m0, 0, p.M, klen, tid);
        lb.load(B, p.b_off, p.b_rs, p.b_cs, n0, 0, p.N, klen, tid);
        la.template store<LDA>(sA[0], tid);
        lb.template store<LDB>(sB[0], tid);
    }
    __syncthreads();

    for (int64_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            la.load(A, p.a_off, p.a_rs, p.a_cs, m0, (kt + 1) * BK, p.M, klen, tid);
            lb.load(B, p.b_off, p.b_rs, p.b_cs, n0, (kt + 1) * BK, p.N, klen, tid);
        }
        const T *a_s = sA[cur];
        const T *b_s = sB[cur];
#pragma unroll
        for (int k = 0; k < BK; k += MF::TK) {
            T af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
                af[i] = a_s[(k + MF::kk(lane)) * LDA + wm * WTM + i * MF::TM + MF::mn(lane)];
#pragma unroll
            for (int j = 0; j < FN; ++j)
                bf[j] = b_s[(k + MF::kk(lane)) * LDB + wn * WTN + j * MF::TN + MF::mn(lane)];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(af[i], bf[j], acc[i][j]);
        }
        if (more) {
            la.template store<LDA>(sA[cur ^ 1], tid);
            lb.template store<LDB>(sB[cur ^ 1], tid);
        }
        __syncthreads();
    }

    /* epilogue: C = alpha*acc + beta*C (C not read when beta == 0) */
    const T alpha = (T)p.alpha, beta = (T)p.beta;
    const bool use_beta = p.beta != 0.0;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int64_t gj = n0 + wn * WTN + j * MF::TN + MF::col(lane);
#pragma unroll
            for (int r = 0; r < MF::NACC; ++r) {
                const int64_t gi = m0 + wm * WTM + i * MF::TM + MF::row(lane, r);
                if (gi < p.M && gj < p.N && (!p.lower || gi >= gj)) {
                    T *cp = C + p.c_off + gi * p.c_rs + gj * p.c_cs;
                    T val = alpha * acc[i][j][r];
                    if (use_beta) val += beta * (*cp);
                    *cp = val;
                    if (p.sh_hi) {
                        const int64_t so = p.sh_off + gi * p.sh_rs + gj * p.sh_cs;
                        const unsigned short hb = bf16_rne((float)val);
                        p.sh_hi[so] = hb;
                        if (p.sh_lo) p.sh_lo[so] = bf16_rne((float)val - __uint_as_float((unsigned)hb << 16));
                    }
                }
            }
        }
}

/* C(i, j) = alpha * sum_z P_z(i, j) + beta * C(i, j) over the split-K partials
 * P_z (column-major M x N, leading dimension M, stride M*N); a lower output
 * touches (and its partials hold) only i >= j. */
template <typename T>
__global__ void splitk_reduce_kernel(const T *P, int64_t M, int64_t N, int64_t splits,
                                     T *C, RfpView vc, double alpha, double beta, int lower)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    const int64_t i = idx % M, j = idx / M;
    if (lower && i < j) return;
    T s = 0;
    for (int64_t z = 0; z < splits; ++z) s += P[z * M * N + idx];
    T *cp = C + vc.at(i, j);
    T val = (T)alpha * s;
    if (beta != 0.0) val += (T)beta * (*cp);
    *cp = val;
}

/* ------------------------------------------------------------------------ */
/* Host launchers                                                            */
/* ------------------------------------------------------------------------ */

template <typename T, typename AS, typename BS, int BM, int BN, int WAVES_M, int BK = 16>
static hipError_t launch_cfg(const GemmArgs &g, hipStream_t s)
{
    KParams p;
    p.a = g.a; p.a_off = g.va.off; p.a_rs = g.va.rs; p.a_cs = g.va.cs;
    p.b = g.b; p.b_off = g.vb.off; p.b_rs = g.vb.rs; p.b_cs = g.vb.cs;
    p.c = g.c; p.c_off = g.vc.off; p.c_rs = g.vc.rs; p.c_cs = g.vc.cs;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.alpha = g.alpha; p.beta = g.beta;
    p.lower = g.lower ? 1 : 0;
    p.tiles_m = (g.M + BM - 1) / BM;
    p.tiles_n = (g.N + BN - 1) / BN;
    p.kchunk = g.K;
    p.c_zstride = 0;
    p.sh_hi = g.c_sh_hi; p.sh_lo = g.c_sh_lo;
    p.sh_off = g.vcs.off; p.sh_rs = g.vcs.rs; p.sh_cs = g.vcs.cs;
    p.batch = g.batch > 1 ? g.batch : 1;
    p.a_bs = g.a_bs; p.b_bs = g.b_bs; p.c_bs = g.c_bs; p.m_last = g.m_last;
    if (p.batch > 1 && (!g.lower || g.splits > 1 || g.inplace || g.c_sh_hi || g.m_last > g.M || g.m_last < 1))
        return hipErrorInvalidValue;
    if (g.lower) {
        if (BM != BN || g.M != g.N) return hipErrorInvalidValue;
        p.ntiles = p.tiles_m * (p.tiles_m + 1) / 2;
    } else {
        p.ntiles = p.tiles_m * p.tiles_n;
    }
    if (g.inplace == 1 && p.tiles_n != 1) return hipErrorInvalidValue;
    if (g.inplace == 2 && p.tiles_m != 1) return hipErrorInvalidValue;
    if (p.ntiles <= 0) return hipSuccess;
    if (p.ntiles > 0x7fffffffLL) return hipErrorInvalidValue;
    /* K == 0 never reads A/B; any unit-stride layout will do */
    const bool akc = g.va.cs == 1;
    const bool bkc = g.vb.cs == 1;
    if (g.K > 0 && ((!akc && g.va.rs != 1) || (!bkc && g.vb.rs != 1))) return hipErrorInvalidValue;

    int64_t splits = 1;
    if (g.splits > 1 && g.partial && !g.inplace) {
        /* split-K: each slice writes a column-major M x N partial; a second
         * kernel sums them into C with alpha / beta */
        splits = g.splits;
        p.kchunk = (g.K + splits - 1) / splits;
        p.kchunk = (p.kchunk + BK - 1) / BK * BK;
        splits = (g.K + p.kchunk - 1) / p.kchunk;
        if (splits < 1) splits = 1;
        p.c = g.partial; p.c_off = 0; p.c_rs = 1; p.c_cs = g.M;
        p.c_zstride = g.M * g.N;
        p.alpha = 1.0; p.beta = 0.0;
        p.sh_hi = p.sh_lo = nullptr;
    }
    dim3 grid((unsigned)p.ntiles, (unsigned)splits, (unsigned)p.batch), block(256);
    if (akc && bkc)  hipLaunchKernelGGL((view_gemm_kernel<T, AS, BS, BM, BN, WAVES_M, BK, true, true>), grid, block, 0, s, p);
    else if (akc)    hipLaunchKernelGGL((view_gemm_kernel<T, AS, BS, BM, BN, WAVES_M, BK, true, false>), grid, block, 0, s, p);
    else if (bkc)    hipLaunchKernelGGL((view_gemm_kernel<T, AS, BS, BM, BN, WAVES_M, BK, false, true>), grid, block, 0, s, p);
    else             hipLaunchKernelGGL((view_gemm_kernel<T, AS, BS, BM, BN, WAVES_M, BK, false, false>), grid, block, 0, s, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || splits == 1 || p.c != g.partial) return e;
    const int64_t total = g.M * g.N;
    hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       (const T *)g.partial, g.M, g.N, splits, (T *)g.c, g.vc, g.alpha, g.beta, p.lower);
    return hipGetLastError();
}

int gemm_inplace_limit(bool f64) { (void)f64; return 128; }

hipError_t launch_gemm_bf16(const GemmArgs &g, RfpMath math, hipStream_t s);

/* Number of 128 x 128 output tiles of a GEMM (on or below the diagonal for a
 * lower-triangular output). */
static int64_t tiles128(const GemmArgs &g)
{
    const int64_t tm = (g.M + 127) / 128, tn = (g.N + 127) / 128;
    return (g.lower ? tm * (tm + 1) / 2 : tm * tn) * (g.batch > 1 ? g.batch : 1);
}

/*
 * Tile choice.  A workgroup's time on a small problem is its own latency:
 * its share of MFMA work plus one exposed global-load round trip per k-slice.
 * Problems that give fewer than SMALL_TILES 128 x 128 tiles (most of the
 * launches deep in the factorisation's recursion) therefore use 64 x 64
 * tiles with BK = 32 — four times the workgroups, a quarter of the work
 * each, half the k-slices — and the in-place TRSM-by-inverse (C aliases A,
 * N <= 128) uses row tiles spanning all N.
 */
constexpr int64_t SMALL_TILES = 128;

hipError_t launch_gemm_f32(const GemmArgs &g, RfpMath math, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    if (math != RFP_MATH_F32 || g.a_bf16) return launch_gemm_bf16(g, math, s);
    /* in-place solves: 64-row tiles, each wave 64 x 32 (32 x 32 tiles with all four waves on one row
     * band: 0.6 ms slower over the 256 tall-panel leaf solves of a BF16 factorisation at n = 65536) */
    if (g.inplace == 1) return launch_cfg<float, float, float, 64, 128, 1, 32>(g, s);
    const bool small = g.inplace != 2 && tiles128(g) < SMALL_TILES;
    if (small) return launch_cfg<float, float, float, 64, 64, 2, 32>(g, s);
    /* narrow outputs (the solves' nrhs columns) use a 128 x 64 tile */
    if (!g.lower && g.N <= 64) return launch_cfg<float, float, float, 128, 64, 4>(g, s);
    return launch_cfg<float, float, float, 128, 128, 2>(g, s);
}

hipError_t launch_gemm_f64(const GemmArgs &g, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    if (g.a_f32) {
        /* FP32-stored operands, FP64 arithmetic (generator residual / norms) */
        if (g.inplace) return hipErrorInvalidValue;
        if (g.lower) {
            /* the Gram matrix op(A)^T op(A): square tiles on or below the diagonal */
            if (!g.b_f32) return hipErrorInvalidValue;
            return launch_cfg<double, float, float, 128, 128, 2>(g, s);
        }
        if (g.b_f32) return launch_cfg<double, float, float, 128, 64, 4>(g, s);
        return launch_cfg<double, float, double, 128, 64, 4>(g, s);
    }
    if (g.inplace == 1) return launch_cfg<double, double, double, 32, 128, 1, 32>(g, s);
    const bool small = g.inplace != 2 && tiles128(g) < SMALL_TILES;
    if (small) return launch_cfg<double, double, double, 64, 64, 2, 32>(g, s);
    if (!g.lower && g.N <= 64) return launch_cfg<double, double, double, 128, 64, 4>(g, s);
    return launch_cfg<double, double, double, 128, 128, 2>(g, s);
}
