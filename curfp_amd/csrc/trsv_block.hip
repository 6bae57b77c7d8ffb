/*
 * trsv_block.hip — triangular solve of one diagonal block (up to 512 rows)
 * against a few right-hand sides, the leaf of spftrs' recursive block
 * substitution.
 *
 * The reference solves with cublasStrsm over whole RFP sub-blocks
 * (src/curfp_spftrs.cpp:258-299).  Here spftrs recurses on the RFP views
 * down to 512-row diagonal blocks (the off-diagonal updates are view GEMMs)
 * and each block is solved by this kernel: one workgroup per 16 right-hand
 * side columns keeps its 512 x 16 slice of B in LDS and walks the block's
 * 128 x 128 leaves in order — forward:  Y_q = inv(L_qq) B_q, then
 * B_r -= L_rq Y_q for the leaves r below; backward (L^T X = B) in reverse
 * with transposed blocks — each product a v_mfma_f32_16x16x4_f32 sequence
 * over a 128 x 128 operand staged in LDS.  The leaf inverses come from the
 * batched inversion kernel (trtri_tiles).
 */
#include "rfp_internal.h"

namespace {

typedef float f32x4v __attribute__((ext_vector_type(4)));

constexpr int LF = 128;        /* leaf */
constexpr int MAXQ = 4;        /* leaves per block: 512 rows */
constexpr int CW = 16;         /* right-hand-side columns per workgroup */
constexpr int AP = LF + 4;     /* LDS pitch of the staged 128 x 128 operand */
constexpr int BP = CW + 1;     /* LDS pitch of the B slice */

/*
 * Stage the 128 x 128 operand A(i, k) = base[off + i*rs + k*cs] (rows /
 * columns at or beyond `valid` read as zero) into As[i*AP + k], reading along
 * the contiguous direction of the source.
 */
__device__ __forceinline__ void stage(float *As, const float *base, int64_t off, int64_t rs, int64_t cs, int valid_i,
                                      int valid_k, int tid)
{
    constexpr int PER = LF * LF / 256;   /* 64 elements per thread, two rows / columns apart */
    float v[PER];
    /* all loads in flight before the first LDS store */
    if (cs == 1) {
        const int k = tid % LF, i0 = tid / LF;
        const float *p = base + off + (int64_t)i0 * rs + k;
        const int64_t step = 2 * rs;
        const bool kv = k < valid_k;
#pragma unroll
        for (int e = 0; e < PER; ++e) v[e] = (kv && i0 + 2 * e < valid_i) ? p[e * step] : 0.f;
#pragma unroll
        for (int e = 0; e < PER; ++e) As[(i0 + 2 * e) * AP + k] = v[e];
    } else {
        const int i = tid % LF, k0 = tid / LF;
        const float *p = base + off + i + (int64_t)k0 * cs;
        const int64_t step = 2 * cs;
        const bool iv = i < valid_i;
#pragma unroll
        for (int e = 0; e < PER; ++e) v[e] = (iv && k0 + 2 * e < valid_k) ? p[e * step] : 0.f;
#pragma unroll
        for (int e = 0; e < PER; ++e) As[i * AP + k0 + 2 * e] = v[e];
    }
}

/*
 * acc (rows wave*32 .. +32, all 16 columns, as two 16 x 16 tiles) =
 * As (128 x 128) * Bs[q0 .. q0+128][0 .. 16)
 */
__device__ __forceinline__ void mma_leaf(f32x4v (&acc)[2], const float *As, const float *Bs, int q0, int wave,
                                         int lane)
{
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t][r] = 0.f;
#pragma unroll 4
    for (int k0 = 0; k0 < LF; k0 += 4) {
        const float b = Bs[(q0 + k0 + lk) * BP + li];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const float a = As[(wave * 32 + t * 16 + li) * AP + k0 + lk];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
        }
    }
}

/* Bs[q0 + row][col] = alpha*acc (+ Bs when add) for the wave's 32 rows */
__device__ __forceinline__ void put_leaf(const f32x4v (&acc)[2], float *Bs, int q0, float alpha, bool add, int wave,
                                         int lane)
{
    const int col = lane & 15;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = q0 + wave * 32 + t * 16 + 4 * (lane >> 4) + r;
            float *p = Bs + row * BP + col;
            *p = add ? *p + alpha * acc[t][r] : alpha * acc[t][r];
        }
}

template <bool TRANS>
__global__ __launch_bounds__(256) void block_solve_kernel(const float *Lbase, RfpView Lv, int m, const float *inv,
                                                          float *Bbase, RfpView Bv, int nrhs)
{
    __shared__ float As[LF * AP];
    __shared__ float Bs[MAXQ * LF * BP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.x * CW;
    const int nq = (m + LF - 1) / LF;
    const int cols = min(CW, nrhs - c0);

    /* B slice: rows [0, m), columns [c0, c0 + cols) */
    {
        constexpr int PER = MAXQ * LF * CW / 256;
        float v[PER];
        const bool col_major = Bv.rs == 1;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = tid + 256 * e;
            const int i = col_major ? idx % (MAXQ * LF) : idx / CW;
            const int c = col_major ? idx / (MAXQ * LF) : idx % CW;
            v[e] = (i < m && c < cols) ? Bbase[Bv.off + (int64_t)i * Bv.rs + (int64_t)(c0 + c) * Bv.cs] : 0.f;
        }
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = tid + 256 * e;
            const int i = col_major ? idx % (MAXQ * LF) : idx / CW;
            const int c = col_major ? idx / (MAXQ * LF) : idx % CW;
            Bs[i * BP + c] = v[e];
        }
    }
    __syncthreads();

    for (int s = 0; s < nq; ++s) {
        const int q = TRANS ? nq - 1 - s : s;
        const int mq = min(LF, m - q * LF);
        /* Y_q = inv(L_qq) B_q   (backward: inv(L_qq)^T B_q) */
        const float *W = inv + (int64_t)q * LF * LF;
        if (TRANS) stage(As, W, 0, 1, LF, mq, mq, tid);
        else       stage(As, W, 0, LF, 1, mq, mq, tid);
        __syncthreads();
        f32x4v acc[2];
        mma_leaf(acc, As, Bs, q * LF, wave, lane);
        __syncthreads();
        put_leaf(acc, Bs, q * LF, 1.f, false, wave, lane);
        __syncthreads();
        /* the other leaves: forward r > q: B_r -= L_rq Y_q; backward r < q: B_r -= L_qr^T X_q */
        const int r_lo = TRANS ? 0 : q + 1, r_hi = TRANS ? q : nq;
        for (int r = r_lo; r < r_hi; ++r) {
            const int mr = min(LF, m - r * LF);
            if (TRANS) {
                /* A(i, k) = L(q*LF + k, r*LF + i) */
                const RfpView blk = Lv.sub(q * LF, r * LF);
                stage(As, Lbase, blk.off, blk.cs, blk.rs, mr, mq, tid);
            } else {
                const RfpView blk = Lv.sub(r * LF, q * LF);
                stage(As, Lbase, blk.off, blk.rs, blk.cs, mr, mq, tid);
            }
            __syncthreads();
            mma_leaf(acc, As, Bs, q * LF, wave, lane);
            __syncthreads();
            put_leaf(acc, Bs, r * LF, -1.f, true, wave, lane);
            __syncthreads();
        }
    }

    for (int idx = tid; idx < MAXQ * LF * CW; idx += 256) {
        int i, c;
        if (Bv.rs == 1) { i = idx % (MAXQ * LF); c = idx / (MAXQ * LF); }
        else            { c = idx % CW; i = idx / CW; }
        if (i < m && c < cols) Bbase[Bv.off + (int64_t)i * Bv.rs + (int64_t)(c0 + c) * Bv.cs] = Bs[i * BP + c];
    }
}

}  // namespace

int block_solve_rows() { return MAXQ * LF; }

hipError_t launch_block_solve_f32(const float *Lbase, RfpView Lv, int m, const float *inv, float *Bbase, RfpView Bv,
                                  int nrhs, bool trans, hipStream_t s)
{
    if (m <= 0 || nrhs <= 0) return hipSuccess;
    if (m > MAXQ * LF) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((nrhs + CW - 1) / CW);
    if (trans)
        hipLaunchKernelGGL((block_solve_kernel<true>), dim3(grid), dim3(256), 0, s, Lbase, Lv, m, inv, Bbase, Bv, nrhs);
    else
        hipLaunchKernelGGL((block_solve_kernel<false>), dim3(grid), dim3(256), 0, s, Lbase, Lv, m, inv, Bbase, Bv,
                           nrhs);
    return hipGetLastError();
}
