/*
 * gemm_bf16_dma.hip — the large BF16 updates of ssfrk (BF16 input) and of
 * spftrf's SYRK / GEMM / TRSM updates in BF16 mode:
 *
 *   C(i, j) = alpha * sum_k A(i, k) * B(j, k) + beta * C(i, j)
 *
 * A and B are BF16 row blocks with k contiguous and 16-byte aligned rows
 * (the factor's BF16 shadow, or the caller's BF16 matrix); C is an FP32 view
 * into the RFP array.
 *
 * Structure: 256 x 256 output tile (128 x 128 for mid-size problems), 4
 * wave64s (2 x 2) of 128 x 128 (64 x 64) each — 4 x 4 (2 x 2) blocks of
 * v_mfma_f32_32x32x16_bf16, 256 (64) FP32 accumulators per lane.
 * k-slices of BK = 64 travel global -> LDS by buffer_load_dwordx4 ... lds
 * (no VGPR staging, no LDS store instructions): one wave instruction fills
 * 1 KB of LDS (8 rows of 128 bytes, whole cache lines) with lane l's 16
 * bytes at offset 16 l, and each lane chooses which global 16-byte chunk it
 * fetches so that LDS chunk p of row r holds global chunk p ^ ((r >> 1) & 7).
 * A 128-byte row spans half of the 64 LDS banks, so with that XOR swizzle
 * the 16 lanes of one LDS cycle (16 consecutive rows, one chunk each) hit
 * all 64 banks once.  Per-tile buffer descriptors bound the rows (num_records covers
 * exactly the valid rows) and chunks at or beyond K get an offset outside
 * the descriptor, so the edges of the problem read as zeros without
 * branches.  Two LDS buffers (128 KB in all at 256 x 256) alternate with the
 * loop unrolled by two, so each buffer is a distinct LDS object and the
 * compiler does not wait on the next slice's loads before reading the
 * current one.
 */
#include "rfp_internal.h"

namespace {

typedef float  f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
/* 16-byte vector that only assumes element alignment (RFP leading dimensions
 * leave columns of a view unaligned) */
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
#define LDS_PTR(p) ((__attribute__((address_space(3))) void *)(p))

constexpr int BK = 64;               /* k per slice */
constexpr int ROWB = BK * 2;         /* bytes of one row of a slice: 8 chunks of 16 */

/* Tile BT x BT (256 or 128): 4 waves of BT/2 x BT/2, F x F blocks of 32 x 32
 * each; an operand slice is BT rows of ROWB bytes, loaded by 2 F wave
 * instructions per wave (8 rows each). */
template <int BT> struct Tile {
    static constexpr int F = BT / 64;
    static constexpr int SLICE = BT * ROWB;
};
constexpr int RSRC_FLAGS = 0x00020000;

struct DmaParams {
    const unsigned short *a; int64_t a_ld;   /* element (i, k) at a[i*a_ld + k] */
    const unsigned short *b; int64_t b_ld;
    float *c; int64_t c_off, c_rs, c_cs;
    int64_t M, N, K;
    float alpha, beta;
    int64_t tiles_m, tiles_n, ntiles;
    int lower;
    int batch;                      /* diagonal blocks along blockIdx.y (GemmArgs::batch) */
    int64_t a_bs, b_bs, c_bs, m_last;
};

__device__ __forceinline__ int64_t xcd_remap(int64_t pid, int64_t total)
{
    const int64_t X = 8;
    const int64_t q = total / X, r = total % X;
    const int64_t x = pid % X, w = pid / X;
    return x * q + (x < r ? x : r) + w;
}

__device__ __forceinline__ void tile_coords(const DmaParams &p, int64_t t, int64_t &tm, int64_t &tn)
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

/* Loads of one operand's k-slice [k0, k0 + BK): groups of 8 rows, q-th of a
 * wave = group 4 q + wave.  Lane l handles row 8 g + l / 8 and LDS chunk
 * l % 8, which holds global chunk (l % 8) ^ ((row >> 1) & 7). */
struct SliceDma {
    int rowoff;   /* (l / 8) * ld2: byte offset of the lane's row within a group */
    int c8;       /* 8 * global chunk of the lane */

    __device__ __forceinline__ void init(int ld2, int lane, int wave)
    {
        const int rl = lane >> 3;
        /* row 8 g + rl with g = 4 q + wave: (row >> 1) & 7 == 4 (wave & 1) + rl / 2 */
        c8 = 8 * ((lane & 7) ^ ((4 * (wave & 1) + (rl >> 1)) & 7));
        rowoff = rl * ld2;
    }
    __device__ __forceinline__ void issue(unsigned char *lds, __amdgpu_buffer_rsrc_t rs, int ld2, int k0, int K,
                                          int q, int wave) const
    {
        const int g = q * 4 + wave;
        const int kk = k0 + c8;
        /* K % 8 == 0: a chunk is wholly inside or outside [0, K) */
        const int off = kk < K ? 8 * g * ld2 + rowoff + kk * 2 : (int)0x80000000;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, LDS_PTR(lds + g * 1024), 16, off, 0, 0, 0);
    }
};

/* Operand fragments of one 16-wide k step t of a slice: lane l reads row
 * l % 32 of each 32-row block, global chunk 2 t + l / 32 (swizzled). */
template <int F> struct Frags {
    bf16x8 a[F], b[F];
};

__device__ __forceinline__ int frag_off(int t, int lane)
{
    /* row = 32 i + (lane & 31) (+ BT/2 wm): (row >> 1) & 7 == ((lane & 31) >> 1) & 7 */
    return (lane & 31) * ROWB + ((((2 * t) + (lane >> 5)) ^ (((lane & 31) >> 1) & 7)) << 4);
}

/*
 * acc += the wave's block of the slice in (As, Bs): four 16-wide k steps,
 * operand fragments double-buffered in registers (step t + 1's read two at a
 * time between step t's groups of F MFMAs), and during step 0 the loads of
 * the next slice (into Na, Nb) issued the same way.  The issue order is fixed
 * with scheduling barriers.  SWAP: the MFMA operands exchanged, so acc[i][j]
 * holds the transpose of C's block (i, j).
 */
template <int BT, bool SWAP>
__device__ __forceinline__ void mma_slice(f32x16 (&acc)[Tile<BT>::F][Tile<BT>::F], const unsigned char *As,
                                          const unsigned char *Bs, unsigned char *Na, unsigned char *Nb,
                                          const SliceDma &da, const SliceDma &db, __amdgpu_buffer_rsrc_t rsa,
                                          __amdgpu_buffer_rsrc_t rsb, int lda2, int ldb2, int k_next, int K,
                                          int wm, int wn, int wave, int lane)
{
    constexpr int F = Tile<BT>::F;
    const unsigned char *arow = As + wm * (BT / 2) * ROWB;
    const unsigned char *brow = Bs + wn * (BT / 2) * ROWB;
    Frags<F> f[2];
    {
        const int o = frag_off(0, lane);
#pragma unroll
        for (int i = 0; i < F; ++i) {
            f[0].a[i] = *(const bf16x8 *)(arow + i * 32 * ROWB + o);
            f[0].b[i] = *(const bf16x8 *)(brow + i * 32 * ROWB + o);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int cur = t & 1, nxt = cur ^ 1;
        const int o = frag_off(t + 1, lane);
#pragma unroll
        for (int i = 0; i < F; ++i) {
            if (t < 3) {
                f[nxt].a[i] = *(const bf16x8 *)(arow + i * 32 * ROWB + o);
                f[nxt].b[i] = *(const bf16x8 *)(brow + i * 32 * ROWB + o);
            }
            if (t < 2) {
                /* the next slice's 2 F groups per operand, spread over steps 0 and 1 */
                da.issue(Na, rsa, lda2, k_next, K, t * F + i, wave);
                db.issue(Nb, rsb, ldb2, k_next, K, t * F + i, wave);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < F; ++j)
                acc[i][j] = SWAP ? __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[cur].b[j], f[cur].a[i], acc[i][j], 0, 0, 0)
                                 : __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[cur].a[i], f[cur].b[j], acc[i][j], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

template <int BT, bool SWAP>
__global__ __launch_bounds__(256) void gemm_bf16_dma_kernel(DmaParams p)
{
    constexpr int F = Tile<BT>::F, SLICE = Tile<BT>::SLICE;
    /* one LDS object per buffer (A slice, then B slice) */
    __shared__ __attribute__((aligned(1024))) unsigned char s0[2 * SLICE];
    __shared__ __attribute__((aligned(1024))) unsigned char s1[2 * SLICE];
    unsigned char *const sA0 = s0, *const sB0 = s0 + SLICE;
    unsigned char *const sA1 = s1, *const sB1 = s1 + SLICE;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    if (p.batch > 1) {
        /* diagonal block bz of a batched lower-triangular launch */
        const int64_t bz = blockIdx.y;
        p.a += bz * p.a_bs;
        p.b += bz * p.b_bs;
        p.c_off += bz * p.c_bs;
        if (bz == p.batch - 1) p.M = p.N = p.m_last;
    }
    int64_t tm, tn;
    tile_coords(p, xcd_remap(blockIdx.x, p.ntiles), tm, tn);
    /* the tile is workgroup-uniform: keep it (and the buffer descriptors
     * built from it) in scalar registers */
    const int64_t m0 = (int64_t)__builtin_amdgcn_readfirstlane((int)tm) * BT;
    const int64_t n0 = (int64_t)__builtin_amdgcn_readfirstlane((int)tn) * BT;
    if (m0 >= p.M || n0 >= p.N) return;   /* past a smaller last block (whole workgroup) */
    const int64_t ra = (p.M - m0) < BT ? (p.M - m0) : BT;
    const int64_t rb = (p.N - n0) < BT ? (p.N - n0) : BT;
    const int lda2 = (int)(p.a_ld * 2), ldb2 = (int)(p.b_ld * 2);
    const __amdgpu_buffer_rsrc_t rsa =
        __builtin_amdgcn_make_buffer_rsrc((void *)(p.a + m0 * p.a_ld), 0, (int)(ra * lda2), RSRC_FLAGS);
    const __amdgpu_buffer_rsrc_t rsb =
        __builtin_amdgcn_make_buffer_rsrc((void *)(p.b + n0 * p.b_ld), 0, (int)(rb * ldb2), RSRC_FLAGS);
    const int K = (int)p.K;

    f32x16 acc[F][F];
#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    SliceDma da, db;
    da.init(lda2, lane, wave);
    db.init(ldb2, lane, wave);
    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
#pragma unroll
        for (int q = 0; q < 2 * F; ++q) {
            da.issue(sA0, rsa, lda2, 0, K, q, wave);
            db.issue(sB0, rsb, ldb2, 0, K, q, wave);
        }
    }
    /* Slice s sits in buffer s % 2; its step 0 issues the loads of slice
     * s + 1 into the other buffer (a slice at or beyond K loads as zeros;
     * an odd slice count multiplies one zero slice, so the body unrolled by
     * two has no data-dependent exit).  Before each slice the wave waits for
     * its own loads, and the barrier makes the buffer complete for all waves
     * and the other buffer free (every LDS read of the previous slice has
     * been consumed by an MFMA before a wave gets there).  A plain
     * s_barrier: nothing else is outstanding. */
    for (int s = 0; s < nk; s += 2) {
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_s_barrier();
        mma_slice<BT, SWAP>(acc, sA0, sB0, sA1, sB1, da, db, rsa, rsb, lda2, ldb2, (s + 1) * BK, K, wm, wn, wave, lane);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_s_barrier();
        mma_slice<BT, SWAP>(acc, sA1, sB1, sA0, sB0, da, db, rsa, rsb, lda2, ldb2, (s + 2) * BK, K, wm, wn, wave, lane);
    }
    /* drain the zero-fill loads of the slice beyond K */
    __builtin_amdgcn_s_waitcnt(0);

    /* Epilogue.  A lane holds, for each 32 x 32 block, one column and rows
     * 4 g + 0..3 (g = 0..3, offset 4 for lanes 32..63): with C a column view
     * (rows contiguous) each such run of 4 rows is one 16-byte store (and
     * load when beta != 0) instead of four scattered ones. */
    const bool use_beta = p.beta != 0.f;
    if (SWAP) {
        /* acc[i][j] is C's block (i, j) transposed: a lane holds row
         * 32 i + lane % 32 and, for each g, 4 consecutive columns 8 g + 4 (lane / 32)
         * .. + 3 — one 16-byte store (and load when beta != 0) into a row view
         * (columns contiguous, c_cs == 1) */
#pragma unroll
        for (int i = 0; i < F; ++i)
#pragma unroll
            for (int j = 0; j < F; ++j) {
                const int64_t gi = m0 + wm * (BT / 2) + i * 32 + (lane & 31);
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int64_t gj0 = n0 + wn * (BT / 2) + j * 32 + 8 * g + 4 * (lane >> 5);
                    if (gi < p.M && gj0 + 3 < p.N && (!p.lower || gi >= gj0 + 3)) {
                        float *cp = p.c + p.c_off + gi * p.c_rs + gj0;
                        f32x4u v = {p.alpha * acc[i][j][4 * g], p.alpha * acc[i][j][4 * g + 1],
                                    p.alpha * acc[i][j][4 * g + 2], p.alpha * acc[i][j][4 * g + 3]};
                        if (use_beta) {
                            const f32x4u o = *(const f32x4u *)cp;
                            v = v + p.beta * o;
                        }
                        *(f32x4u *)cp = v;
                        continue;
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int64_t gj = gj0 + e;
                        if (gi < p.M && gj < p.N && (!p.lower || gi >= gj)) {
                            float *cp = p.c + p.c_off + gi * p.c_rs + gj;
                            float val = p.alpha * acc[i][j][4 * g + e];
                            if (use_beta) val += p.beta * (*cp);
                            *cp = val;
                        }
                    }
                }
            }
        return;
    }
    const bool rows_contig = p.c_rs == 1;
#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j) {
            const int64_t gj = n0 + wn * (BT / 2) + j * 32 + (lane & 31);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int64_t gi0 = m0 + wm * (BT / 2) + i * 32 + 8 * g + 4 * (lane >> 5);
                if (rows_contig && gi0 + 3 < p.M && gj < p.N && (!p.lower || gi0 >= gj)) {
                    float *cp = p.c + p.c_off + gi0 + gj * p.c_cs;
                    f32x4u v = {p.alpha * acc[i][j][4 * g], p.alpha * acc[i][j][4 * g + 1],
                                p.alpha * acc[i][j][4 * g + 2], p.alpha * acc[i][j][4 * g + 3]};
                    if (use_beta) {
                        const f32x4u o = *(const f32x4u *)cp;
                        v = v + p.beta * o;
                    }
                    *(f32x4u *)cp = v;
                    continue;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t gi = gi0 + e;
                    if (gi < p.M && gj < p.N && (!p.lower || gi >= gj)) {
                        float *cp = p.c + p.c_off + gi * p.c_rs + gj * p.c_cs;
                        float val = p.alpha * acc[i][j][4 * g + e];
                        if (use_beta) val += p.beta * (*cp);
                        *cp = val;
                    }
                }
            }
        }
}

}  // namespace

/*
 * Launch the DMA kernel for a BF16 row-block GEMM, or return
 * hipErrorNotSupported when the operands do not meet its requirements
 * (16-byte aligned rows and bases, K a multiple of 8, row offsets of a tile
 * within 32 bits); the caller then uses the register-staged kernels.
 */
hipError_t launch_gemm_bf16_dma(const GemmArgs &g, int bt, hipStream_t s)
{
    if (bt != 128 && bt != 256) return hipErrorNotSupported;
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    if (!g.bf16_rows && !g.a_bf16) return hipErrorNotSupported;
    if (g.va.cs != 1 || g.vb.cs != 1) return hipErrorNotSupported;
    if ((g.va.rs % 8) || (g.vb.rs % 8) || (g.va.off % 8) || (g.vb.off % 8) || (g.K % 8)) return hipErrorNotSupported;
    if (((uintptr_t)g.a % 16) || ((uintptr_t)g.b % 16)) return hipErrorNotSupported;
    if (g.va.rs < g.K || g.vb.rs < g.K) return hipErrorNotSupported;
    const int64_t lim = (int64_t)1 << 30;
    if ((int64_t)bt * g.va.rs * 2 >= lim || (int64_t)bt * g.vb.rs * 2 >= lim || g.K >= lim)
        return hipErrorNotSupported;
    if (g.lower && g.M != g.N) return hipErrorInvalidValue;
    DmaParams p;
    p.a = (const unsigned short *)g.a + g.va.off; p.a_ld = g.va.rs;
    p.b = (const unsigned short *)g.b + g.vb.off; p.b_ld = g.vb.rs;
    p.c = (float *)g.c; p.c_off = g.vc.off; p.c_rs = g.vc.rs; p.c_cs = g.vc.cs;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.alpha = (float)g.alpha; p.beta = (float)g.beta;
    p.lower = g.lower ? 1 : 0;
    p.batch = g.batch > 1 ? g.batch : 1;
    p.a_bs = g.a_bs; p.b_bs = g.b_bs; p.c_bs = g.c_bs; p.m_last = g.m_last;
    if (p.batch > 1 && (!g.lower || g.m_last > g.M || g.m_last < 1)) return hipErrorInvalidValue;
    /* a later block's row offsets must stay within the 32-bit descriptors too */
    if (p.batch > 1 && ((p.a_bs % 8) || (p.b_bs % 8))) return hipErrorNotSupported;
    p.tiles_m = (g.M + bt - 1) / bt;
    p.tiles_n = (g.N + bt - 1) / bt;
    p.ntiles = g.lower ? p.tiles_m * (p.tiles_m + 1) / 2 : p.tiles_m * p.tiles_n;
    if (p.ntiles <= 0) return hipSuccess;
    if (p.ntiles > 0x7fffffffLL) return hipErrorInvalidValue;
    /* C a row view (columns contiguous) read and written (beta != 0, the
     * factorisation's updates): the transposed accumulator layout gives the
     * epilogue 16-byte loads and stores along rows.  With beta == 0 the
     * plain layout's 4-byte stores already cover whole 128-byte lines of a
     * row per instruction and measured faster (ssfrk). */
    const bool swap = g.vc.cs == 1 && g.vc.rs != 1 && g.beta != 0.0;
    const dim3 grid((unsigned)p.ntiles, (unsigned)p.batch);
    if (bt == 256) {
        if (swap) hipLaunchKernelGGL((gemm_bf16_dma_kernel<256, true>), grid, dim3(256), 0, s, p);
        else      hipLaunchKernelGGL((gemm_bf16_dma_kernel<256, false>), grid, dim3(256), 0, s, p);
    } else {
        if (swap) hipLaunchKernelGGL((gemm_bf16_dma_kernel<128, true>), grid, dim3(256), 0, s, p);
        else      hipLaunchKernelGGL((gemm_bf16_dma_kernel<128, false>), grid, dim3(256), 0, s, p);
    }
    return hipGetLastError();
}
