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
constexpr int AP = LF + 4;     /* LDS pitch of a staged operand held as As[i][k] */
constexpr int ATP = LF + 16;   /* LDS pitch of a staged operand held as As[k][i] */
constexpr int MAXSTEPS = MAXQ + MAXQ * (MAXQ - 1) / 2;
constexpr int BP = CW + 1;     /* LDS pitch of the B slice */

/* The 128 x 128 operand A(i, k) = base[off + i*rs + k*cs] of one step (rows /
 * columns at or beyond vi / vk read as zero). */
struct Blk {
    const float *base;
    int64_t off, rs, cs;
    int vi, vk;
};

/* 16-byte vector that only assumes element alignment (RFP leading dimensions
 * such as n + 1 leave rows unaligned) */
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

/* A thread's 64 elements of a staged operand as 16 vectors of 4 consecutive
 * elements along the contiguous direction of the source (eight rows /
 * columns apart). */
struct Staged {
    f32x4v v[16];

    __device__ __forceinline__ void load(const Blk &b, int tid)
    {
        const int c4 = (tid % 32) * 4, o0 = tid / 32;
        const bool kc = b.cs == 1;
        /* kc: lanes run along k (c4 = k), o = i; else along i (c4 = i), o = k */
        const int vc = kc ? b.vk : b.vi, vo = kc ? b.vi : b.vk;
        const int64_t so = kc ? b.rs : b.cs;
        const float *p = b.base + b.off + c4 + (int64_t)o0 * so;
        const int64_t step = 8 * so;
        const bool full = c4 + 4 <= vc;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int o = o0 + 8 * e;
            const float *q = p + e * step;
            if (o < vo && full) {
                const f32x4u x = *(const f32x4u *)q;
                v[e] = f32x4v{x[0], x[1], x[2], x[3]};
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) v[e][c] = (o < vo && c4 + c < vc) ? q[c] : 0.f;
            }
        }
    }
    /* into LDS along the source's contiguous direction: As[i*AP + k] when k
     * is contiguous (kc), else As[k*ATP + i] (both pitches keep the 16-byte
     * stores aligned and the MFMA operand reads free of bank conflicts) */
    __device__ __forceinline__ void store(float *As, bool kc, int tid) const
    {
        const int c4 = (tid % 32) * 4, o0 = tid / 32;
        const int pitch = kc ? AP : ATP;
#pragma unroll
        for (int e = 0; e < 16; ++e) *(f32x4v *)(As + (o0 + 8 * e) * pitch + c4) = v[e];
    }
};

/*
 * acc (rows wave*32 .. +32, all 16 columns, as two 16 x 16 tiles) =
 * As (128 x 128) * Bs[q0 .. q0+128][0 .. 16)
 */
__device__ __forceinline__ void mma_leaf(f32x4v (&acc)[2], const float *As, bool kc, const float *Bs, int q0,
                                         int wave, int lane)
{
    const int li = lane & 15, lk = lane >> 4;
    /* element (i, k) of the operand at As[i*ri + k*rk] */
    const int ri = kc ? AP : 1, rk = kc ? 1 : ATP;
    const float *a0 = As + (wave * 32 + li) * ri + lk * rk;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t][r] = 0.f;
#pragma unroll
    for (int k0 = 0; k0 < LF; k0 += 4) {
        const float b = Bs[(q0 + k0 + lk) * BP + li];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const float a = a0[t * 16 * ri + k0 * rk];
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
    __shared__ __attribute__((aligned(16))) float As[LF * ATP];
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

    /*
     * The block's steps in order — forward: for each leaf q, Y_q = inv(L_qq) B_q,
     * then B_r -= L_rq Y_q for r > q; backward: for q from the last leaf,
     * X_q = inv(L_qq)^T B_q, then B_r -= L_qr^T X_q for r < q.
     */
    auto step_of = [&](int j, int &q, int &r) {
        /* j -> (q, r): r == q for the leaf inverse, else the update of leaf r */
        for (int s = 0; s < nq; ++s) {
            const int qq = TRANS ? nq - 1 - s : s;
            const int nr = TRANS ? qq : nq - 1 - qq;   /* updates after this leaf */
            if (j <= nr) {
                q = qq;
                r = j == 0 ? qq : (TRANS ? j - 1 : qq + j);
                return;
            }
            j -= nr + 1;
        }
        q = r = 0;
    };
    auto blk_of = [&](int q, int r) {
        const int mq = min(LF, m - q * LF), mr = min(LF, m - r * LF);
        if (r == q) {
            const float *W = inv + (int64_t)q * LF * LF;
            return TRANS ? Blk{W, 0, 1, LF, mq, mq} : Blk{W, 0, LF, 1, mq, mq};
        }
        if (TRANS) {
            /* A(i, k) = L(q*LF + k, r*LF + i) */
            const RfpView v = Lv.sub(q * LF, r * LF);
            return Blk{Lbase, v.off, v.cs, v.rs, mr, mq};
        }
        const RfpView v = Lv.sub(r * LF, q * LF);
        return Blk{Lbase, v.off, v.rs, v.cs, mr, mq};
    };
    const int nsteps = nq + nq * (nq - 1) / 2;
    /* two register sets: set j % 2 holds step j's operand; steps j + 1 and
     * j + 2 are in flight while step j's MFMAs run */
    Staged st[2];
    Blk bk[2];
    int qs[2], rs[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (j < nsteps) {
            step_of(j, qs[j], rs[j]);
            bk[j] = blk_of(qs[j], rs[j]);
            st[j].load(bk[j], tid);
        }
#pragma unroll
    for (int j = 0; j < MAXSTEPS; ++j) {
        if (j < nsteps) {
            const int x = j & 1;
            const bool kc = bk[x].cs == 1;
            const int cq = qs[x], cr = rs[x];
            st[x].store(As, kc, tid);
            __syncthreads();
            if (j + 2 < nsteps) {
                step_of(j + 2, qs[x], rs[x]);
                bk[x] = blk_of(qs[x], rs[x]);
                st[x].load(bk[x], tid);
            }
            f32x4v acc[2];
            mma_leaf(acc, As, kc, Bs, cq * LF, wave, lane);
            __syncthreads();
            if (cr == cq) put_leaf(acc, Bs, cq * LF, 1.f, false, wave, lane);
            else          put_leaf(acc, Bs, cr * LF, -1.f, true, wave, lane);
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
