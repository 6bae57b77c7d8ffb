/*
 * trsv_block.hip — spftrs' triangular substitutions.
 *
 * The reference solves with cublasStrsm over whole RFP sub-blocks
 * (src/curfp_spftrs.cpp:258-299).  Two kernels here:
 *
 * sweep_solve_kernel (FP32, nrhs <= 4096): one launch per direction over the
 * whole factor.  The factor is cut into 128-row blocks (v11's, then v22's);
 * one workgroup per block and group of <= 64 right-hand sides accumulates
 * the products of its row of L tiles with the already solved blocks, then
 * applies the diagonal tile's inverse.  Workgroups publish finished blocks
 * through a progress word and wait on it before using a block, so the
 * substitution's serial chain runs at the speed of one tile product per
 * block instead of one kernel launch per recursion step.
 *
 * block_solve_kernel (used by the recursive substitution for more right-hand
 * sides): one workgroup per 16 right-hand-side columns keeps its 512 x 16
 * slice of B in LDS and walks the block's 128 x 128 leaves in order —
 * forward:  Y_q = inv(L_qq) B_q, then B_r -= L_rq Y_q for the leaves r below;
 * backward (L^T X = B) in reverse with transposed blocks — each product a
 * v_mfma_f32_16x16x4_f32 sequence over a 128 x 128 operand staged in LDS.
 *
 * The leaf inverses come from the batched inversion kernel (trtri_tiles).
 *
 * sfmm_kernel (ssfmm, FP32) reuses the sweep's block geometry and products:
 * C = alpha * A * Y + beta * C with A the symmetric RFP matrix, one
 * workgroup per 128-row block of A and group of <= 64 columns of Y, summing
 * A(b, c) Y_c over every block c with no dependencies between workgroups.
 * sfmm64_kernel is its FP64 twin (VALU FMAs, 64 x 64 output tiles).
 */
#include "rfp_internal.h"
#include "tile_stage.h"

namespace {

constexpr int LF = TILE_LF;    /* leaf */
constexpr int MAXQ = 4;        /* leaves per block: 512 rows */
constexpr int CW = 16;         /* right-hand-side columns per workgroup */
constexpr int AP = LF + 4;     /* LDS pitch of a staged operand held as As[i][k] */
constexpr int ATP = LF + 16;   /* LDS pitch of a staged operand held as As[k][i] */
constexpr int MAXSTEPS = MAXQ + MAXQ * (MAXQ - 1) / 2;
constexpr int BP = CW + 1;     /* LDS pitch of the B slice */

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
            st[j].template load<false>(bk[j], tid);
        }
#pragma unroll
    for (int j = 0; j < MAXSTEPS; ++j) {
        if (j < nsteps) {
            const int x = j & 1;
            const bool kc = bk[x].cs == 1;
            const int cq = qs[x], cr = rs[x];
            st[x].store(As, kc ? AP : ATP, tid);
            __syncthreads();
            if (j + 2 < nsteps) {
                step_of(j + 2, qs[x], rs[x]);
                bk[x] = blk_of(qs[x], rs[x]);
                st[x].template load<false>(bk[x], tid);
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


/* ------------------------------------------------------------------------ */
/* FP64 block solve                                                          */
/* ------------------------------------------------------------------------ */
/*
 * The FP64 twin of block_solve_kernel: one workgroup per 16 right-hand-side
 * columns keeps its (up to) 512 x 16 slice of B in LDS and walks the block's
 * 64-row leaves in the same order — forward, for each leaf q, Y_q =
 * inv(L_qq) B_q then B_r -= L_rq Y_q (r > q); backward, from the last leaf,
 * X_q = inv(L_qq)^T B_q then B_r -= L_qr^T X_q (r < q) — each step one
 * 64 x 64 operand (a leaf inverse or an off-diagonal tile of L, transposed
 * for the backward solve) times the slice's 64 x 16 leaf on
 * v_mfma_f64_16x16x4_f64.
 *
 * Wave w owns output rows 16 w .. 16 w + 15 of every leaf: its lanes load
 * their own operand values straight from memory into the MFMA's A layout
 * (no LDS staging), D steps ahead, and an update step writes only the
 * wave's own rows of B_r while reading Y_q, which no update changes.  So the
 * waves synchronise only around the inverse steps (all updates of B_q done
 * before Y_q is formed from all of it, and Y_q complete before it is used),
 * three barriers per leaf instead of two per step.  The k chunks alternate
 * between two accumulators, so consecutive MFMAs are independent, and the
 * zero half of a triangular inverse is skipped.
 */
namespace solve64 {

constexpr int LF = 64;                   /* FP64 leaf */
constexpr int MQ = 8;                    /* leaves per block: 512 rows */
constexpr int CW = 16;                   /* right-hand-side columns per workgroup */
constexpr int BP = CW + 1;               /* LDS pitch of the B slice (doubles) */
constexpr int BPER = MQ * LF * CW / 256; /* B-slice elements per thread */
constexpr int KC = LF / 4;               /* k chunks of one operand row */
constexpr int D = 4;                     /* steps of operand loads in flight */

typedef double f64x4v __attribute__((ext_vector_type(4)));

/* operand A(i, k) = base[off + i*rs + k*cs] for i < vi, k < vk (zero beyond) */
struct Op {
    const double *base;
    int64_t off, rs, cs;
    int vi, vk;
};

/* element e of thread tid of the B slice as (row i, column c) */
__device__ __forceinline__ void slice_elem(bool col_major, int tid, int e, int &i, int &c)
{
    const int idx = tid + 256 * e;
    if (col_major) { i = idx % (MQ * LF); c = idx / (MQ * LF); }
    else           { c = idx % CW;        i = idx / CW; }
}

template <bool TRANS>
__global__ __launch_bounds__(256) void block_solve64_kernel(const double *Lbase, RfpView Lv, int m, const double *inv,
                                                            double *Bbase, RfpView Bv, int nrhs)
{
    __shared__ double Bs[MQ * LF * BP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.x * CW;
    const int nq = (m + LF - 1) / LF;
    const int cols = min(CW, nrhs - c0);
    const bool col_major = Bv.rs == 1;

    /* B slice: rows [0, m), columns [c0, c0 + cols); zero beyond */
    {
        double v[BPER];
#pragma unroll
        for (int e = 0; e < BPER; ++e) {
            int i, c;
            slice_elem(col_major, tid, e, i, c);
            v[e] = (i < m && c < cols) ? Bbase[Bv.off + (int64_t)i * Bv.rs + (int64_t)(c0 + c) * Bv.cs] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < BPER; ++e) {
            int i, c;
            slice_elem(col_major, tid, e, i, c);
            Bs[i * BP + c] = v[e];
        }
    }

    auto op_of = [&](int q, int r) {
        const int mq = min(LF, m - q * LF), mr = min(LF, m - r * LF);
        if (r == q) {
            const double *W = inv + (int64_t)q * LF * LF;
            return TRANS ? Op{W, 0, 1, LF, mq, mq} : Op{W, 0, LF, 1, mq, mq};
        }
        if (TRANS) {
            /* A(i, k) = L(q*LF + k, r*LF + i) */
            const RfpView v = Lv.sub((int64_t)q * LF, (int64_t)r * LF);
            return Op{Lbase, v.off, v.cs, v.rs, mr, mq};
        }
        const RfpView v = Lv.sub((int64_t)r * LF, (int64_t)q * LF);
        return Op{Lbase, v.off, v.rs, v.cs, mr, mq};
    };

    /* The lane's A values of a step: row ar = 16 w + lane % 16 of the operand,
     * k = 4 c + lane / 16 for chunk c.  Every load is unconditional and used
     * as loaded — an element outside the operand reads an exact zero of the
     * first leaf inverse (row 0, column 63: above its diagonal), and the steps
     * past the last re-load the last step's operand — so that each step waits
     * for its own operand only. */
    const int li = lane & 15, lk = lane >> 4;
    const int ar = wave * 16 + li;
    const double *const zero = inv + (LF - 1);
    double x[D][KC];
    int qs[D], rs[D];
    auto load = [&](int d, const Op &o) {
        const double *p = o.base + o.off + (int64_t)ar * o.rs + (int64_t)lk * o.cs;
        const int64_t s4 = 4 * o.cs;
        if (o.vi == LF && o.vk == LF) {
            /* whole operand (all but a ragged last leaf): no masks */
#pragma unroll
            for (int c = 0; c < KC; ++c) x[d][c] = p[c * s4];
            return;
        }
        const bool iok = ar < o.vi;
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const double *ptr = (iok && 4 * c + lk < o.vk) ? p + c * s4 : zero;
            x[d][c] = *ptr;
        }
    };
    /* The step after (q, r) in solve order (the last step repeats).  Forward:
     * for each leaf q, the inverse (q, q), then the updates r = q + 1 ..;
     * backward: from the last leaf, (q, q), then r = 0 .. q - 1. */
    auto next = [&](int &q, int &r) {
        if (TRANS) {
            if (r == q) { if (q > 0) r = 0; }
            else if (r + 1 < q) ++r;
            else { --q; r = q; }
        } else {
            if (r + 1 < nq) ++r;
            else if (q + 1 < nq) { ++q; r = q; }
        }
    };

    const int nsteps = nq + nq * (nq - 1) / 2;
    /* (pq, pr): the step whose operand is loaded next (the first: the inverse
     * of the first leaf in solve order) */
    int pq = TRANS ? nq - 1 : 0, pr = pq;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        qs[d] = pq; rs[d] = pr;
        load(d, op_of(pq, pr));
        next(pq, pr);
    }
    for (int j0 = 0; j0 < nsteps; j0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int j = j0 + d;
            if (j >= nsteps) break;
            const int cq = qs[d], cr = rs[d];
            const bool inv_step = cq == cr;
            if (inv_step) lds_barrier();   /* every wave's updates of B_q are in */
            /* acc = A (rows 16 w ..) * Bs[leaf cq] (64 x 16); a triangular
             * inverse's zero 16-column chunks skipped (wave-uniform) */
            int klo = 0, khi = LF;
            if (inv_step) {
                if (TRANS) klo = 16 * wave;
                else       khi = 16 * (wave + 1);
            }
            const double *bcol = Bs + (cq * LF + lk) * BP + li;
            f64x4v acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < KC; c += 2) {
                if (4 * c < klo || 4 * c >= khi) continue;
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x[d][c], bcol[4 * c * BP], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x[d][c + 1], bcol[4 * (c + 1) * BP], acc1, 0, 0, 0);
            }
            qs[d] = pq; rs[d] = pr;
            load(d, op_of(pq, pr));
            next(pq, pr);
            /* output row 16 w + lk + 4 rr of leaf cr, column li (the FP64
             * MFMA layout): the wave's own rows */
            double *dst = Bs + (cr * LF + wave * 16 + lk) * BP + li;
            if (inv_step) {
                lds_barrier();             /* every wave has read all of B_q */
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) dst[4 * rr * BP] = acc0[rr] + acc1[rr];
                lds_barrier();             /* Y_q complete */
            } else {
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) dst[4 * rr * BP] -= acc0[rr] + acc1[rr];
            }
        }
    }
    lds_barrier();

    for (int e = 0; e < BPER; ++e) {
        int i, c;
        slice_elem(col_major, tid, e, i, c);
        if (i < m && c < cols) Bbase[Bv.off + (int64_t)i * Bv.rs + (int64_t)(c0 + c) * Bv.cs] = Bs[i * BP + c];
    }
}

}  // namespace solve64

/* ------------------------------------------------------------------------ */
/* Whole-factor sweep: one launch per substitution direction                 */
/* ------------------------------------------------------------------------ */

/*
 * LDS layouts of the sweep.  The MFMA k index of lane group lk (= lane / 16)
 * in the s-th product of a 16-wide k chunk is 4 lk + s, so a lane's four
 * operand values of a chunk are consecutive in k:
 *   L tile, k contiguous in HBM:  As[i*SP + k]    (one 16-byte read per chunk)
 *   L tile, i contiguous in HBM:  As[k*SP + i]    (four 4-byte reads)
 *   Y block:                      Ys[col*SP + k]  (one 16-byte read)
 * SP = 132 keeps the 16-byte stores in HBM order and all of these reads free
 * of bank conflicts.  A row-major B is transposed into Ys by 4-byte stores.
 */
constexpr int SP = LF + 4;
__host__ __device__ constexpr int ys_elems(int cw) { return cw * SP; }

struct SweepArgs {
    const float *L;
    RfpView v11, v21, v22;
    int64_t n1, n2;
    int s1, nb;                   /* 128-row blocks in v11, in all */
    const float *inv;             /* inverse of block b's diagonal tile at inv + b*128*128 (row-major) */
    const float *pm;              /* P_b,l = inv(L_bb) L_b,b-l and inv(L_bb)^T L_b+l,b^T (l = 1, 2)
                                   * at pm + ((2*dir + l - 1)*nb + b)*128*128 (row-major) */
    float *B;
    RfpView bv;                   /* B(i, c) = B[bv.off + i*bv.rs + c*bv.cs] */
    int nrhs, groups;
    unsigned long long *prog;     /* [groups]: (epoch << 32) | blocks finished, in solve order */
    unsigned *tick;               /* [0] ticket counter, [1] finished workgroups */
    unsigned long long epoch;
    int *err;
    int wait_limit;               /* polls per wait before giving up */
};

/* global first row, view-local first row and row count of block b */
__device__ __forceinline__ void sweep_block(const SweepArgs &a, int b, int64_t &grow, int64_t &lrow, int &rows)
{
    if (b < a.s1) {
        lrow = (int64_t)b * LF;
        grow = lrow;
        rows = (int)min<int64_t>(LF, a.n1 - lrow);
    } else {
        lrow = (int64_t)(b - a.s1) * LF;
        grow = a.n1 + lrow;
        rows = (int)min<int64_t>(LF, a.n2 - lrow);
    }
}

/* block (r, c), r > c, of L as a view */
__device__ __forceinline__ RfpView sweep_tile(const SweepArgs &a, int r, int c)
{
    int64_t gr, lr, gc, lc;
    int mr, mc;
    sweep_block(a, r, gr, lr, mr);
    sweep_block(a, c, gc, lc, mc);
    if (r < a.s1) return a.v11.sub(lr, lc);
    if (c < a.s1) return a.v21.sub(lr, lc);
    return a.v22.sub(lr, lc);
}

/* A thread's share of one Y block (128 x CW) moved between B and LDS as
 * 16-byte vectors along B's contiguous direction. */
template <int CW, bool YCM>
struct YStage {
    static constexpr int PER = LF * CW / 4 / 256;
    f32x4v v[PER];

    /* element e of the thread: (row, col) of the block, 4 consecutive along
     * rows (column-major B) or along columns (row-major B) */
    __device__ __forceinline__ static void pos(int tid, int e, int &row, int &col)
    {
        if (YCM) {
            row = (tid % 32) * 4;
            col = tid / 32 + 8 * e;
        } else {
            constexpr int TPR = CW / 4;   /* threads per row */
            col = (tid % TPR) * 4;
            row = tid / TPR + (256 / TPR) * e;
        }
    }
    __device__ __forceinline__ static float *lds(float *Ys, int row, int col) { return Ys + col * SP + row; }
    __device__ __forceinline__ void load(const SweepArgs &a, int64_t grow, int rows, int col0, int cols, int tid)
    {
        const int64_t sr = a.bv.rs, sc = a.bv.cs;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            int row, col;
            pos(tid, e, row, col);
            const float *p = a.B + a.bv.off + (grow + row) * sr + (int64_t)(col0 + col) * sc;
            const bool full = YCM ? (row + 4 <= rows && col < cols) : (row < rows && col + 4 <= cols);
            if (full) {
                const f32x4u x = *(const f32x4u *)p;
                v[e] = f32x4v{x[0], x[1], x[2], x[3]};
            } else {
                const int64_t st = YCM ? sr : sc;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool ok = YCM ? (row + q < rows && col < cols) : (row < rows && col + q < cols);
                    v[e][q] = ok ? p[q * st] : 0.f;
                }
            }
        }
    }
    __device__ __forceinline__ void store_lds(float *Ys, int tid) const
    {
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            int row, col;
            pos(tid, e, row, col);
            if (YCM) {
                *(f32x4v *)lds(Ys, row, col) = v[e];
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) *lds(Ys, row, col + q) = v[e][q];
            }
        }
    }
    __device__ __forceinline__ void load_lds(float *Ys, int tid)
    {
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            int row, col;
            pos(tid, e, row, col);
            if (YCM) {
                v[e] = *(const f32x4v *)lds(Ys, row, col);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) v[e][q] = *lds(Ys, row, col + q);
            }
        }
    }
    __device__ __forceinline__ void store(const SweepArgs &a, int64_t grow, int rows, int col0, int cols, int tid) const
    {
        const int64_t sr = a.bv.rs, sc = a.bv.cs;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            int row, col;
            pos(tid, e, row, col);
            float *p = a.B + a.bv.off + (grow + row) * sr + (int64_t)(col0 + col) * sc;
            const bool full = YCM ? (row + 4 <= rows && col < cols) : (row < rows && col + 4 <= cols);
            if (full) {
                *(f32x4u *)p = f32x4u{v[e][0], v[e][1], v[e][2], v[e][3]};
            } else {
                const int64_t st = YCM ? sr : sc;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool ok = YCM ? (row + q < rows && col < cols) : (row < rows && col + q < cols);
                    if (ok) p[q * st] = v[e][q];
                }
            }
        }
    }
};

/* The wave's two 16-row tiles of a 128-row block: w and 7 - w, so that the
 * triangular diagonal solve gives every wave the same number of k chunks. */
__device__ __forceinline__ int sweep_row_tile(int wave, int t) { return t == 0 ? wave : 7 - wave; }

/* Operands of one 16-wide k chunk: the lane's 4 consecutive k of each of its
 * two A rows and CT B columns. */
template <int CT> struct ChunkOps {
    f32x4v a[2], b[CT];
};

template <int CT, bool AKC>
__device__ __forceinline__ void chunk_load(ChunkOps<CT> &o, const float *As, const float *Ys, int q, int wave,
                                           int lane)
{
    const int li = lane & 15, k = q * 16 + 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int i = sweep_row_tile(wave, t) * 16 + li;
        if (AKC) {
            o.a[t] = *(const f32x4v *)(As + i * SP + k);
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) o.a[t][s] = As[(k + s) * SP + i];
        }
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) o.b[ct] = *(const f32x4v *)(Ys + (ct * 16 + li) * SP + k);
}

/* acc[t][ct] (row tile sweep_row_tile(wave, t), columns 16 ct ..) += A (128 x
 * 128) * Y, the next chunk's operands read while the current chunk's MFMAs
 * run.  TRI = 1 / 2: A is lower / upper triangular by 16 x 16 tiles (the
 * diagonal block's inverse), and the k chunks that only meet zero tiles of
 * a row tile are skipped. */
template <int CT, bool AKC, int TRI>
__device__ __forceinline__ void mma_sweep(f32x4v (&acc)[2][CT], const float *As, const float *Ys, int wave, int lane)
{
    auto need = [&](int t, int q) {
        const int rt = sweep_row_tile(wave, t);
        return TRI == 0 || (TRI == 1 ? q <= rt : q >= rt);
    };
    ChunkOps<CT> o[2];
    chunk_load<CT, AKC>(o[0], As, Ys, 0, wave, lane);
#pragma unroll
    for (int q = 0; q < LF / 16; ++q) {
        if (q + 1 < LF / 16) chunk_load<CT, AKC>(o[(q + 1) & 1], As, Ys, q + 1, wave, lane);
        const ChunkOps<CT> &c = o[q & 1];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (!need(t, q)) continue;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
                    acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(c.a[t][s], c.b[ct][s], acc[t][ct], 0, 0, 0);
        }
    }
}

typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8v round_bf16(const f32x4v &lo, const f32x4v &hi)
{
    return bf16x8v{(__bf16)lo[0], (__bf16)lo[1], (__bf16)lo[2], (__bf16)lo[3],
                   (__bf16)hi[0], (__bf16)hi[1], (__bf16)hi[2], (__bf16)hi[3]};   /* nearest even */
}

/*
 * acc += A (128 x 128) * Y with both operands rounded to BF16 and
 * v_mfma_f32_16x16x32_bf16 (FP32 accumulation), reading the same FP32 LDS
 * layouts as mma_sweep.  A 32-wide k chunk p is the FP32 layout's 16-wide
 * chunks 2p and 2p + 1: lane group lk supplies k = 32 p + 4 lk + s and
 * 32 p + 16 + 4 lk + s (s < 4) for both operands, so the pairing of A's and
 * Y's k is the same and the LDS reads stay those of mma_sweep (conflict
 * free).  The next chunk's FP32 operands are read before the current
 * chunk's MFMAs and rounded after them.
 */
template <int CT, bool AKC>
__device__ __forceinline__ void mma_sweep_bf16(f32x4v (&acc)[2][CT], const float *As, const float *Ys, int wave,
                                               int lane)
{
    bf16x8v a[2], b[CT];
    ChunkOps<CT> lo, hi;
    chunk_load<CT, AKC>(lo, As, Ys, 0, wave, lane);
    chunk_load<CT, AKC>(hi, As, Ys, 1, wave, lane);
#pragma unroll
    for (int p = 0; p < LF / 32; ++p) {
#pragma unroll
        for (int t = 0; t < 2; ++t) a[t] = round_bf16(lo.a[t], hi.a[t]);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) b[ct] = round_bf16(lo.b[ct], hi.b[ct]);
        if (p + 1 < LF / 32) {
            chunk_load<CT, AKC>(lo, As, Ys, 2 * p + 2, wave, lane);
            chunk_load<CT, AKC>(hi, As, Ys, 2 * p + 3, wave, lane);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[ct], acc[t][ct], 0, 0, 0);
    }
}

/*
 * The products P_b,l that take the diagonal inverse off the sweep's serial
 * chain (one workgroup per block b, direction and l = 1, 2):
 *   forward:   P_b,l = inv(L_bb) L_b,b-l,       Y_b = inv(L_bb) R_b - P_b,2 Y_b-2 - P_b,1 Y_b-1
 *   backward:  P_b,l = inv(L_bb)^T L_b+l,b^T,   X_b = inv(L_bb)^T R_b - P_b,2 X_b+2 - P_b,1 X_b+1
 * with R_b the right-hand side less every other contribution.  They depend
 * only on the factor (128^3 products, once per solve).
 */
__global__ __launch_bounds__(256) void sweep_pmat_kernel(SweepArgs a)
{
    __shared__ __attribute__((aligned(16))) float Ia[LF * SP];
    __shared__ __attribute__((aligned(16))) float Lt[LF * SP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.x, dir = blockIdx.y >> 1, lag = 1 + (blockIdx.y & 1);
    const int c = dir == 0 ? b - lag : b + lag;
    if (c < 0 || c >= a.nb) return;
    int64_t gb, lb, gc, lc;
    int mb, mc;
    sweep_block(a, b, gb, lb, mb);
    sweep_block(a, c, gc, lc, mc);
    {
        Staged st;
        st.load(Blk{a.inv + (int64_t)b * LF * LF, 0, LF, 1, mb, mb}, tid);
        st.store(Ia, SP, tid);                                   /* Ia[i*SP + k] = inv(i, k) */
    }
    /* the off-diagonal tile in its natural orientation: L(r, q) at Lt[r*lr + q*lq] */
    const RfpView v = dir == 0 ? sweep_tile(a, b, c) : sweep_tile(a, c, b);
    const bool kc = v.cs == 1;
    {
        Staged st;
        st.load(Blk{a.L, v.off, v.rs, v.cs, dir == 0 ? mb : mc, dir == 0 ? mc : mb}, tid);
        st.store(Lt, SP, tid);
    }
    const int lr = kc ? SP : 1, lq = kc ? 1 : SP;
    __syncthreads();
    /* P(i, j) = sum_k A(i, k) B(k, j): A = inv or inv^T, B(k, j) = L_b,c(k, j) or L_c,b(j, k) */
    const int ai = dir == 0 ? SP : 1, ak = dir == 0 ? 1 : SP;
    const int bk = dir == 0 ? lr : lq, bj = dir == 0 ? lq : lr;
    const int li = lane & 15, lk = lane >> 4;
    f32x4v acc[2][8];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < 8; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[t][ct][r] = 0.f;
    for (int k0 = 0; k0 < LF; k0 += 4) {
        const int k = k0 + lk;
        float av[2], bv[8];
#pragma unroll
        for (int t = 0; t < 2; ++t) av[t] = Ia[(wave * 32 + t * 16 + li) * ai + k * ak];
#pragma unroll
        for (int ct = 0; ct < 8; ++ct) bv[ct] = Lt[k * bk + (ct * 16 + li) * bj];
#pragma unroll
        for (int ct = 0; ct < 8; ++ct)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t], bv[ct], acc[t][ct], 0, 0, 0);
    }
    float *P = (float *)a.pm + ((int64_t)(2 * dir + lag - 1) * a.nb + b) * LF * LF;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < 8; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wave * 32 + t * 16 + 4 * lk + r, j = ct * 16 + li;
                P[i * LF + j] = (i < mb && j < mc) ? acc[t][ct][r] : 0.f;
            }
}

/* A workgroup's wait for an earlier block is bounded by a.wait_limit polls
 * (by default 2^20 polls of >= ~250 ns each, a quarter of a second or more,
 * against a few milliseconds for a whole sweep), so that every wave exits
 * even if the progress words were corrupted.  After a timeout the workgroup
 * stops waiting, sets bit 0 of *err and writes its block as NaN. */

/*
 * Forward (L Y = B) or backward (L^T X = Y) substitution over the whole
 * factor in one launch.  Workgroups take tickets in order; ticket t solves
 * block o = t / groups in solve order (block b = o forward, nb - 1 - o
 * backward) for the CW columns of group g = t % groups:
 *
 *   forward:   Y_b = inv(L_bb) (B_b - sum_{c < b} L_bc Y_c)
 *   backward:  X_b = inv(L_bb)^T (Y_b - sum_{c > b} L_cb^T X_c)
 *
 * The contributions are taken in solve order, each after waiting until
 * group g's progress word says block c is finished; then the workgroup
 * publishes its own block.  A workgroup only ever waits for workgroups with
 * smaller tickets, which have already started (tickets are handed out at
 * start), so the sweep completes whatever the number of resident
 * workgroups.  For block o in solve order the steps are: the
 * contributions of blocks 0 .. o-3 (solve order), then Z = inv * (B_b - sum)
 * (kept in LDS), then acc = P_b,2 Y_c(o-2) + P_b,1 Y_c(o-1) and Y_b = Z - acc
 * (fewer look-back steps for o < 2), so that only one tile product separates
 * the previous block's publication from this block's, and the block before
 * that leaves a whole step of slack for the rest.  The operand of step j + 2 (an L tile, the inverse or P_b) is
 * loaded into registers while step j runs, and the next step's Y block (or
 * B_b) as soon as it is known to be finished.
 * BF16: the contributions of blocks 0 .. o-3 and the look-back products
 * are BF16 products (mma_sweep_bf16); the diagonal inverse stays FP32.
 * With column-major B the block is stored straight from the accumulators
 * (each lane's 4 consecutive rows of a column as one 16-byte store).
 */
template <int CW, bool TRANS, bool YCM, bool BF16>
__global__ __launch_bounds__(256) void sweep_solve_kernel(SweepArgs a)
{
    constexpr int CT = CW / 16;
    __shared__ __attribute__((aligned(16))) float As[LF * SP];
    __shared__ __attribute__((aligned(16))) float Ys[ys_elems(CW)];
    __shared__ __attribute__((aligned(16))) float Zs[ys_elems(CW)];
    __shared__ unsigned s_ticket;
    __shared__ int s_seen, s_fail;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    if (tid == 0) {
        s_ticket = __hip_atomic_fetch_add(a.tick, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_seen = 0;
        s_fail = 0;
    }
    __syncthreads();
    /* wave-uniform values in scalar registers */
    const unsigned ticket = __builtin_amdgcn_readfirstlane(s_ticket);
    if (ticket >= gridDim.x) {
        /* the counter was shared with another launch: this block index would
         * lie outside the factor, so touch nothing and report it (bit 1) */
        if (tid == 0) __hip_atomic_fetch_or(a.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const int o = (int)(ticket / (unsigned)a.groups), g = (int)(ticket % (unsigned)a.groups);
    const int b = TRANS ? a.nb - 1 - o : o;
    const int col0 = g * CW, cols = min(CW, a.nrhs - col0);
    int64_t grow, lrow;
    int rows;
    sweep_block(a, b, grow, lrow, rows);

    /* steps: j < zs the contribution of block c(j) (solve order), j == zs the
     * inverse applied to B_b less those contributions, j > zs the product of
     * P_b,l with block c(j - 1), l = o + 1 - j */
    const int npl = o < 2 ? o : 2;
    const int zs = o - npl;
    const int ns = o + 1;
    auto cblk = [&](int j) { return TRANS ? a.nb - 1 - j : j; };
    auto tile_of = [&](int j) {
        if (j == zs) {
            const float *W = a.inv + (int64_t)b * LF * LF;
            return TRANS ? Blk{W, 0, 1, LF, rows, rows} : Blk{W, 0, LF, 1, rows, rows};
        }
        const int c = cblk(j < zs ? j : j - 1);
        int64_t gc, lc;
        int mc;
        sweep_block(a, c, gc, lc, mc);
        if (j > zs) {
            const int l = o + 1 - j;
            const float *P = a.pm + ((int64_t)(2 * (TRANS ? 1 : 0) + l - 1) * a.nb + b) * LF * LF;
            return Blk{P, 0, LF, 1, rows, mc};
        }
        if (TRANS) {
            const RfpView v = sweep_tile(a, c, b);      /* A(i, k) = L(c-row k, b-row i) */
            return Blk{a.L, v.off, v.cs, v.rs, rows, mc};
        }
        const RfpView v = sweep_tile(a, b, c);
        return Blk{a.L, v.off, v.rs, v.cs, rows, mc};
    };
    /* blocks in solve order that must be finished before step j's Y block is read */
    auto need_of = [&](int j) { return j < zs ? j + 1 : (j == zs ? 0 : j); };
    /* wait until the first `need` blocks in solve order are finished */
    auto wait_blocks = [&](int need) {
        if (tid == 0) {
            const unsigned long long want = (a.epoch << 32) | (unsigned long long)need;
            unsigned long long v = 0;
            int it = 0;
            for (;;) {
                v = __hip_atomic_load(a.prog + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (v >= want) break;
                if (++it > a.wait_limit) {
                    __hip_atomic_fetch_or(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_fail = 1;
                    v = 0x7fffffffull;   /* no further waits */
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            s_seen = (int)(v & 0xffffffffu);
        }
        __syncthreads();
    };
    auto load_y = [&](YStage<CW, YCM> &y, int j) {
        int64_t gc = grow, lc;
        int mc = rows;
        if (j != zs) sweep_block(a, cblk(j < zs ? j : j - 1), gc, lc, mc);
        y.load(a, gc, mc, col0, cols, tid);
    };

    f32x4v acc[2][CT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[t][ct][r] = 0.f;

    Staged st[2];
    bool kc0, kc1 = false;
    {
        const Blk bk = tile_of(0);
        kc0 = bk.cs == 1;
        st[0].load(bk, tid);
    }
    if (ns > 1) {
        const Blk bk = tile_of(1);
        kc1 = bk.cs == 1;
        st[1].load(bk, tid);
    }
    YStage<CW, YCM> yv;
    bool yready = false;
    int seen = 0;
    for (int j = 0; j < ns; ++j) {
        const int x = j & 1;
        if (!yready) {
            const int need = need_of(j);
            if (seen < need) {
                wait_blocks(need);
                seen = __builtin_amdgcn_readfirstlane(s_seen);
            }
            load_y(yv, j);
        }
        lds_barrier();     /* the previous step's MMAs are done with As and Ys */
        const bool kc = x ? kc1 : kc0;
        if (x) st[1].store(As, SP, tid);
        else   st[0].store(As, SP, tid);
        yv.store_lds(Ys, tid);
        lds_barrier();
        /* next step's Y block before the operand two steps ahead: the loads
         * complete in issue order, so the next step waits for its Y block
         * without waiting for that operand */
        if (j + 1 < ns) {
            yready = seen >= need_of(j + 1);
            if (yready) load_y(yv, j + 1);
        }
        if (j + 2 < ns) {
            const Blk bk = tile_of(j + 2);
            if (x) { kc1 = bk.cs == 1; st[1].load(bk, tid); }
            else   { kc0 = bk.cs == 1; st[0].load(bk, tid); }
        }
        if (j < zs) {
            if (BF16) {
                if (kc) mma_sweep_bf16<CT, true>(acc, As, Ys, wave, lane);
                else    mma_sweep_bf16<CT, false>(acc, As, Ys, wave, lane);
            } else {
                if (kc) mma_sweep<CT, true, 0>(acc, As, Ys, wave, lane);
                else    mma_sweep<CT, false, 0>(acc, As, Ys, wave, lane);
            }
        } else if (j == zs) {
            /* R = B_b - acc in place in Ys, then acc = inv(L_bb) R (transposed backward) */
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = sweep_row_tile(wave, t) * 16 + 4 * (lane >> 4) + r, col = ct * 16 + (lane & 15);
                        *YStage<CW, YCM>::lds(Ys, row, col) -= acc[t][ct][r];
                        acc[t][ct][r] = 0.f;
                    }
            lds_barrier();
            /* inv(L_bb) is lower triangular, its transpose upper */
            if (TRANS) mma_sweep<CT, false, 2>(acc, As, Ys, wave, lane);
            else       mma_sweep<CT, true, 1>(acc, As, Ys, wave, lane);
            if (o > 0) {
                /* Z = acc waits in LDS for the last block */
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = sweep_row_tile(wave, t) * 16 + 4 * (lane >> 4) + r;
                            const int col = ct * 16 + (lane & 15);
                            *YStage<CW, YCM>::lds(Zs, row, col) = acc[t][ct][r];
                            acc[t][ct][r] = 0.f;
                        }
            }
        } else {
            /* acc += P_b,l Y_c; after the last, the block's result is Z - acc */
            if (BF16) mma_sweep_bf16<CT, true>(acc, As, Ys, wave, lane);
            else      mma_sweep<CT, true, 0>(acc, As, Ys, wave, lane);
            if (j == o) {
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = sweep_row_tile(wave, t) * 16 + 4 * (lane >> 4) + r;
                            const int col = ct * 16 + (lane & 15);
                            acc[t][ct][r] = *YStage<CW, YCM>::lds(Zs, row, col) - acc[t][ct][r];
                        }
            }
        }
    }
    const float nan = __builtin_nanf("");
    const bool fail = s_fail != 0;
    if (YCM) {
        /* rows contiguous in B: a lane's 4 rows of a column in one store */
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int row = sweep_row_tile(wave, t) * 16 + 4 * (lane >> 4), col = ct * 16 + (lane & 15);
                if (col >= cols) continue;
                float *p = a.B + a.bv.off + (grow + row) + (int64_t)(col0 + col) * a.bv.cs;
                const f32x4v v = fail ? f32x4v{nan, nan, nan, nan} : acc[t][ct];
                if (row + 4 <= rows) {
                    *(f32x4u *)p = f32x4u{v[0], v[1], v[2], v[3]};
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (row + r < rows) p[r] = v[r];
                }
            }
    } else {
        lds_barrier();
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = sweep_row_tile(wave, t) * 16 + 4 * (lane >> 4) + r, col = ct * 16 + (lane & 15);
                    *YStage<CW, YCM>::lds(Ys, row, col) = fail ? nan : acc[t][ct][r];
                }
        lds_barrier();
        yv.load_lds(Ys, tid);
        yv.store(a, grow, rows, col0, cols, tid);
    }
    /* publish: every wave waits until its stores have reached L2; then one
     * wave writes the L2 back (agent-scope release fence), waits for that
     * write-back, and moves the progress word — a single write-back for the
     * workgroup instead of one per wave */
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_store(a.prog + g, (a.epoch << 32) | (unsigned long long)(o + 1), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        /* the last workgroup to finish re-arms the ticket counter for the next launch */
        if (__hip_atomic_fetch_add(a.tick + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
            __hip_atomic_store(a.tick, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.tick + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* ssfmm / dsfmm: symmetric RFP matrix times a dense block                   */
/* ------------------------------------------------------------------------ */

struct SfmmArgs {
    SweepArgs s;        /* geometry, A at s.L, Y at s.B / s.bv, its column count in s.nrhs */
    float *C;
    RfpView cv;         /* C(i, c) = C[cv.off + i*cv.rs + c*cv.cs] */
    float alpha, beta;
};

/*
 * C(rows of block b, columns of group y) = alpha * sum_c A(b, c) Y_c + beta C.
 * Tile (b, c) of the symmetric A is the stored tile for b > c, the stored
 * tile (c, b) read transposed for b < c, and for b == c the stored triangle
 * read with mirrored indices (so no entry outside the triangle is read).  The
 * next off-diagonal tile is staged in registers while the current products
 * run; the next Y block is loaded as soon as the current one is in LDS.
 */
template <int CW, bool YCM>
__global__ __launch_bounds__(256) void sfmm_kernel(SfmmArgs p)
{
    constexpr int CT = CW / 16;
    __shared__ __attribute__((aligned(16))) float As[LF * SP];
    __shared__ __attribute__((aligned(16))) float Ys[ys_elems(CW)];
    const SweepArgs &a = p.s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.x;
    const int col0 = blockIdx.y * CW, cols = min(CW, a.nrhs - col0);
    int64_t grow, lrow;
    int rows;
    sweep_block(a, b, grow, lrow, rows);
    auto tile_of = [&](int c) {
        int64_t gc, lc;
        int mc;
        sweep_block(a, c, gc, lc, mc);
        if (b > c) {
            const RfpView v = sweep_tile(a, b, c);
            return Blk{a.L, v.off, v.rs, v.cs, rows, mc};
        }
        const RfpView v = sweep_tile(a, c, b);      /* A(b-row i, c-col k) = stored (c-row k, b-col i) */
        return Blk{a.L, v.off, v.cs, v.rs, rows, mc};
    };
    auto load_y = [&](YStage<CW, YCM> &y, int c) {
        int64_t gc, lc;
        int mc;
        sweep_block(a, c, gc, lc, mc);
        y.load(a, gc, mc, col0, cols, tid);
    };

    f32x4v acc[2][CT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[t][ct][r] = 0.f;

    /* alpha == 0: A and Y are not read, C := beta C */
    const int nsteps = p.alpha == 0.f ? 0 : a.nb;
    Staged st;
    bool kcn = false;
    {
        const int c0 = b == 0 ? 1 : 0;
        if (c0 < nsteps) {
            const Blk bk = tile_of(c0);
            kcn = bk.cs == 1;
            st.load(bk, tid);
        }
    }
    YStage<CW, YCM> yv;
    if (nsteps > 0) load_y(yv, 0);
    for (int c = 0; c < nsteps; ++c) {
        lds_barrier();     /* the previous products are done with As and Ys */
        bool kc = true;
        if (c == b) {
            const RfpView v = sweep_tile(a, b, b);
            for (int e = 0; e < LF * LF / 256; ++e) {
                const int idx = tid + 256 * e;
                const int i = idx / LF, k = idx % LF;
                As[i * SP + k] = (i < rows && k < rows) ? a.L[v.at(max(i, k), min(i, k))] : 0.f;
            }
        } else {
            st.store(As, SP, tid);
            kc = kcn;
            const int cn = c + 1 == b ? c + 2 : c + 1;
            if (cn < nsteps) {
                const Blk bk = tile_of(cn);
                kcn = bk.cs == 1;
                st.load(bk, tid);
            }
        }
        yv.store_lds(Ys, tid);
        lds_barrier();
        if (c + 1 < nsteps) load_y(yv, c + 1);
        if (kc) mma_sweep<CT, true, 0>(acc, As, Ys, wave, lane);
        else    mma_sweep<CT, false, 0>(acc, As, Ys, wave, lane);
    }
    const bool use_beta = p.beta != 0.f;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = sweep_row_tile(wave, t) * 16 + 4 * (lane >> 4) + r, col = ct * 16 + (lane & 15);
                if (row < rows && col < cols) {
                    float *q = p.C + p.cv.off + (grow + row) * p.cv.rs + (int64_t)(col0 + col) * p.cv.cs;
                    const float v = p.alpha * acc[t][ct][r];
                    *q = use_beta ? v + p.beta * *q : v;
                }
            }
}

struct Sfmm64Args {
    const double *A;
    RfpView v11, v21, v22;
    int64_t n, n1;
    const double *B;
    RfpView bv;
    double *C;
    RfpView cv;
    int64_t N;
    double alpha, beta;
};

/* element (r, c) of the symmetric matrix from its stored lower triangle */
__device__ __forceinline__ double sym64(const Sfmm64Args &p, int64_t r, int64_t c)
{
    const int64_t i = r > c ? r : c, j = r > c ? c : r;
    if (i < p.n1) return p.A[p.v11.at(i, j)];
    if (j < p.n1) return p.A[p.v21.at(i - p.n1, j)];
    return p.A[p.v22.at(i - p.n1, j - p.n1)];
}

/* FP64: 64 x 64 tile of C per workgroup, 4 x 4 outputs per thread, k in
 * slices of 16 through LDS */
__global__ __launch_bounds__(256) void sfmm64_kernel(Sfmm64Args p)
{
    constexpr int T = 64, KS = 16;
    __shared__ double As[KS][T + 1];
    __shared__ double Bs[KS][T + 1];
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    const int64_t r0 = (int64_t)blockIdx.x * T, c0 = (int64_t)blockIdx.y * T;
    double acc[4][4] = {};
    const int64_t kend = p.alpha == 0.0 ? 0 : p.n;   /* alpha == 0: A and Y are not read */
    for (int64_t k0 = 0; k0 < kend; k0 += KS) {
        for (int e = tid; e < KS * T; e += 256) {
            const int kk = e / T, x = e % T;
            const int64_t gk = k0 + kk;
            As[kk][x] = (r0 + x < p.n && gk < p.n) ? sym64(p, r0 + x, gk) : 0.0;
            Bs[kk][x] = (gk < p.n && c0 + x < p.N) ? p.B[p.bv.off + gk * p.bv.rs + (c0 + x) * p.bv.cs] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                av[u] = As[kk][tr + 16 * u];
                bv[u] = Bs[kk][tc + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int w = 0; w < 4; ++w) acc[u][w] = fma(av[u], bv[w], acc[u][w]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int64_t r = r0 + tr + 16 * u, c = c0 + tc + 16 * w;
            if (r < p.n && c < p.N) {
                double *q = p.C + p.cv.off + r * p.cv.rs + c * p.cv.cs;
                const double v = p.alpha * acc[u][w];
                *q = p.beta != 0.0 ? v + p.beta * *q : v;
            }
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

int block_solve64_rows() { return solve64::MQ * solve64::LF; }

hipError_t launch_block_solve_f64(const double *Lbase, RfpView Lv, int m, const double *inv, double *Bbase,
                                  RfpView Bv, int nrhs, bool trans, hipStream_t s)
{
    if (m <= 0 || nrhs <= 0) return hipSuccess;
    if (m > block_solve64_rows()) return hipErrorInvalidValue;
    if ((Lv.rs != 1 && Lv.cs != 1) || (Bv.rs != 1 && Bv.cs != 1)) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((nrhs + solve64::CW - 1) / solve64::CW);
    if (trans)
        hipLaunchKernelGGL((solve64::block_solve64_kernel<true>), dim3(grid), dim3(256), 0, s, Lbase, Lv, m, inv,
                           Bbase, Bv, nrhs);
    else
        hipLaunchKernelGGL((solve64::block_solve64_kernel<false>), dim3(grid), dim3(256), 0, s, Lbase, Lv, m, inv,
                           Bbase, Bv, nrhs);
    return hipGetLastError();
}

static SweepArgs sweep_args(const float *L, const RfpGeom &g, const float *inv, const float *pm)
{
    SweepArgs a{};
    a.L = L;
    a.v11 = g.v11; a.v21 = g.v21; a.v22 = g.v22;
    a.n1 = g.n1; a.n2 = g.n2;
    a.s1 = (int)((g.n1 + LF - 1) / LF);
    a.nb = a.s1 + (int)((g.n2 + LF - 1) / LF);
    a.inv = inv;
    a.pm = pm;
    return a;
}

hipError_t launch_sweep_pmat_f32(const float *L, const RfpGeom &g, const float *inv, float *pm, hipStream_t s)
{
    if (g.n <= 0) return hipSuccess;
    const SweepArgs a = sweep_args(L, g, inv, pm);
    hipLaunchKernelGGL(sweep_pmat_kernel, dim3((unsigned)a.nb, 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

/* layout of the SWEEP_SYNC_BYTES area: SWEEP_MAX_GROUPS progress words,
 * the ticket and finish counters, the error word */
int *sweep_err_word(void *sync)
{
    return (int *)((unsigned *)((unsigned long long *)sync + SWEEP_MAX_GROUPS) + 2);
}

hipError_t launch_sweep_solve_f32(const float *L, const RfpGeom &g, const float *inv, const float *pm, float *B,
                                  RfpView bv, int nrhs, bool trans, bool bf16, void *sync,
                                  unsigned long long epoch, int wait_limit, hipStream_t s)
{
    if (g.n <= 0 || nrhs <= 0) return hipSuccess;
    if (wait_limit < 0) return hipErrorInvalidValue;
    const int cw = nrhs <= 16 ? 16 : nrhs <= 32 ? 32 : 64;
    const int groups = (nrhs + cw - 1) / cw;
    if (groups > SWEEP_MAX_GROUPS) return hipErrorInvalidValue;
    SweepArgs a = sweep_args(L, g, inv, pm);
    a.B = B; a.bv = bv;
    a.nrhs = nrhs; a.groups = groups;
    a.prog = (unsigned long long *)sync;
    a.tick = (unsigned *)((unsigned long long *)sync + SWEEP_MAX_GROUPS);
    a.err = sweep_err_word(sync);
    a.epoch = epoch;
    a.wait_limit = wait_limit;
    const unsigned grid = (unsigned)a.nb * (unsigned)groups;
    const bool ycm = bv.rs == 1;
    /* the BF16 products are built for column-major right-hand sides only */
    if (bf16 && !ycm) return hipErrorInvalidValue;
#define SWEEP_LAUNCH(CW_, T_, C_, B_) \
    hipLaunchKernelGGL((sweep_solve_kernel<CW_, T_, C_, B_>), dim3(grid), dim3(256), 0, s, a)
#define SWEEP_LAUNCH_CW(CW_)                                                                         \
    do {                                                                                             \
        if (bf16) { if (trans) SWEEP_LAUNCH(CW_, true, true, true); else SWEEP_LAUNCH(CW_, false, true, true); } \
        else if (trans) { if (ycm) SWEEP_LAUNCH(CW_, true, true, false);  else SWEEP_LAUNCH(CW_, true, false, false); } \
        else            { if (ycm) SWEEP_LAUNCH(CW_, false, true, false); else SWEEP_LAUNCH(CW_, false, false, false); } \
    } while (0)
    if (cw == 16) SWEEP_LAUNCH_CW(16);
    else if (cw == 32) SWEEP_LAUNCH_CW(32);
    else SWEEP_LAUNCH_CW(64);
#undef SWEEP_LAUNCH_CW
#undef SWEEP_LAUNCH
    return hipGetLastError();
}

hipError_t launch_sfmm_f32(const float *A, const RfpGeom &g, const float *B, RfpView bv, float *C, RfpView cv,
                           int64_t N, float alpha, float beta, hipStream_t s)
{
    if (g.n <= 0 || N <= 0) return hipSuccess;
    const int cw = N <= 16 ? 16 : N <= 32 ? 32 : 64;
    const int64_t groups = (N + cw - 1) / cw;
    if (groups > 65535 || N > INT32_MAX) return hipErrorInvalidValue;
    SfmmArgs p{};
    p.s = sweep_args(A, g, nullptr, nullptr);
    p.s.B = const_cast<float *>(B);
    p.s.bv = bv;
    p.s.nrhs = (int)N;
    p.C = C;
    p.cv = cv;
    p.alpha = alpha;
    p.beta = beta;
    const dim3 grid((unsigned)p.s.nb, (unsigned)groups);
    const bool ycm = bv.rs == 1;
#define SFMM_LAUNCH(CW_)                                                                          \
    do {                                                                                          \
        if (ycm) hipLaunchKernelGGL((sfmm_kernel<CW_, true>), grid, dim3(256), 0, s, p);          \
        else     hipLaunchKernelGGL((sfmm_kernel<CW_, false>), grid, dim3(256), 0, s, p);         \
    } while (0)
    if (cw == 16) SFMM_LAUNCH(16);
    else if (cw == 32) SFMM_LAUNCH(32);
    else SFMM_LAUNCH(64);
#undef SFMM_LAUNCH
    return hipGetLastError();
}

hipError_t launch_sfmm_f64(const double *A, const RfpGeom &g, const double *B, RfpView bv, double *C, RfpView cv,
                           int64_t N, double alpha, double beta, hipStream_t s)
{
    if (g.n <= 0 || N <= 0) return hipSuccess;
    const int64_t gx = (g.n + 63) / 64, gy = (N + 63) / 64;
    if (gy > 65535) return hipErrorInvalidValue;
    Sfmm64Args p;
    p.A = A; p.v11 = g.v11; p.v21 = g.v21; p.v22 = g.v22; p.n = g.n; p.n1 = g.n1;
    p.B = B; p.bv = bv; p.C = C; p.cv = cv; p.N = N; p.alpha = alpha; p.beta = beta;
    hipLaunchKernelGGL(sfmm64_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, p);
    return hipGetLastError();
}
