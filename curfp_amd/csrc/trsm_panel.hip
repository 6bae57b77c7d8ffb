/*
 * trsm_panel.hip — the bottom of spftrf's recursive triangular solve: one
 * launch for X := X * L^{-T} with L a lower-triangular diagonal block of at
 * most four 128-row leaves (the reference's TRSM step of the RFP Cholesky,
 * src/curfp_spftrf.cpp:215-236, reached here through the recursion over the
 * three RFP views).
 *
 * The rows of X are independent right-hand sides of the substitution, so
 * each workgroup takes 32 rows of X (a slab, held in LDS) and walks the
 * leaves of L in order:
 *
 *     X_q := (X_q - sum_{p < q} X_p L_qp^T) * inv(L_qq)^T
 *
 * with the leaf inverses written by the leaf Cholesky kernel.  This replaces
 * the recursion's last two levels — four one-leaf in-place GEMMs and three
 * small update GEMMs — by one launch.  Every L tile and inverse passes
 * through registers into LDS, loaded during the previous step's products; the products
 * are v_mfma_f32_16x16x4_f32 over 16-wide k chunks (each lane's four k of a
 * chunk consecutive, so slab and k-contiguous tiles are read 16 bytes at a
 * time).  The final X is written to the RFP array and, in the BF16 modes,
 * rounded into the factor's BF16 shadow (hi and, for split-BF16, lo part).
 */
#include "rfp_internal.h"
#include "tile_stage.h"
#include "trsm_panel_shape.h"

namespace {

using panel_shape::LF;
using panel_shape::MAXQ;
using panel_shape::RM;
using panel_shape::XP;   /* == 4 mod 64: conflict-free 16-byte reads */
using panel_shape::SP;
static_assert(LF == TILE_LF && panel_shape::THREADS == 256, "panel shape");

struct PanelArgs {
    float *a;                  /* RFP array */
    RfpView X;                 /* mb x m */
    RfpView L;                 /* m x m lower */
    const float *inv;          /* inverse of leaf q at inv + q*LF*LF, row-major */
    int mb, m;
    unsigned short *sh_hi, *sh_lo;
    RfpView S;                 /* shadow view of X */
};

__device__ __forceinline__ unsigned short bf16_rne(float x)
{
    __bf16 b = (__bf16)x;
    return __builtin_bit_cast(unsigned short, b);
}

/* The wave's two 16-column tiles of a 128-column leaf: w and 7 - w. */
__device__ __forceinline__ int col_tile(int wave, int c) { return c == 0 ? wave : 7 - wave; }

struct Ops {
    f32x4v a[2], b[2];
};

/* operands of k chunk q16 (k = 16 q16 + 4 (lane / 16) + s for the s-th
 * product): A = slab columns x0 .., B(k, j) = tile(j, k) */
template <bool BKC>
__device__ __forceinline__ void ops_load(Ops &o, const float *Xs, int x0, const float *Ls, int q16, int wave, int lane)
{
    const int li = lane & 15, k = q16 * 16 + 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < 2; ++t) o.a[t] = *(const f32x4v *)(Xs + (t * 16 + li) * XP + x0 + k);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int j = col_tile(wave, c) * 16 + li;
        if (BKC) {
            o.b[c] = *(const f32x4v *)(Ls + j * SP + k);
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) o.b[c][s] = Ls[(k + s) * SP + j];
        }
    }
}

/* acc[t][c] (rows 16 t .., columns col_tile(wave, c) ..) += slab(:, x0 ..
 * x0 + 128) * tile^T.  The inverse tiles are lower triangular; skipping the
 * k chunks past a column tile's diagonal measured slower than multiplying
 * their zeros (0.4 ms per factorisation at n = 65536), so every chunk runs. */
template <bool BKC>
__device__ __forceinline__ void mma_panel(f32x4v (&acc)[2][2], const float *Xs, int x0, const float *Ls, int wave,
                                          int lane)
{
    Ops o[2];
    ops_load<BKC>(o[0], Xs, x0, Ls, 0, wave, lane);
#pragma unroll
    for (int q = 0; q < LF / 16; ++q) {
        if (q + 1 < LF / 16) ops_load<BKC>(o[(q + 1) & 1], Xs, x0, Ls, q + 1, wave, lane);
        const Ops &cur = o[q & 1];
        /* four independent accumulators interleaved */
#pragma unroll
        for (int s = 0; s < 4; ++s) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.a[t][s], cur.b[0][s], acc[t][0], 0, 0, 0);
                acc[t][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.a[t][s], cur.b[1][s], acc[t][1], 0, 0, 0);
            }
        }
    }
}

typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8v round_bf16(const f32x4v &lo, const f32x4v &hi)
{
    return bf16x8v{(__bf16)lo[0], (__bf16)lo[1], (__bf16)lo[2], (__bf16)lo[3],
                   (__bf16)hi[0], (__bf16)hi[1], (__bf16)hi[2], (__bf16)hi[3]};   /* nearest even */
}

/* mma_panel (no triangular skip) with both operands rounded to BF16 and
 * v_mfma_f32_16x16x32_bf16: a 32-wide k chunk p is the FP32 chunks 2p and
 * 2p + 1, lane group lk supplying k = 32 p + 4 lk + s and 32 p + 16 + 4 lk + s
 * (s < 4) for both operands, so the LDS reads are those of mma_panel. */
template <bool BKC>
__device__ __forceinline__ void mma_panel_bf16(f32x4v (&acc)[2][2], const float *Xs, int x0, const float *Ls,
                                               int wave, int lane)
{
    bf16x8v a[2], b[2];
    Ops lo, hi;
    ops_load<BKC>(lo, Xs, x0, Ls, 0, wave, lane);
    ops_load<BKC>(hi, Xs, x0, Ls, 1, wave, lane);
#pragma unroll
    for (int q = 0; q < LF / 32; ++q) {
#pragma unroll
        for (int t = 0; t < 2; ++t) a[t] = round_bf16(lo.a[t], hi.a[t]);
#pragma unroll
        for (int c = 0; c < 2; ++c) b[c] = round_bf16(lo.b[c], hi.b[c]);
        if (q + 1 < LF / 32) {
            ops_load<BKC>(lo, Xs, x0, Ls, 2 * q + 2, wave, lane);
            ops_load<BKC>(hi, Xs, x0, Ls, 2 * q + 3, wave, lane);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c)
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[c], acc[t][c], 0, 0, 0);
    }
}

/* mma_panel with both operands split into BF16 high and low parts, three
 * v_mfma_f32_16x16x32_bf16 per 32-wide k chunk (hi hi^T + hi lo^T + lo hi^T,
 * FP32 accumulation): products to ~2^-16 of FP32's, the lo lo^T term (2^-16
 * relative) dropped.  Same LDS reads as mma_panel_bf16. */
template <bool BKC>
__device__ __forceinline__ void mma_panel_bf16x3(f32x4v (&acc)[2][2], const float *Xs, int x0, const float *Ls,
                                                 int wave, int lane)
{
    Ops lo, hi;
    ops_load<BKC>(lo, Xs, x0, Ls, 0, wave, lane);
    ops_load<BKC>(hi, Xs, x0, Ls, 1, wave, lane);
#pragma unroll
    for (int q = 0; q < LF / 32; ++q) {
        bf16x8v ah[2], al[2], bh[2], bl[2];
        auto split = [](const f32x4v &x0, const f32x4v &x1, bf16x8v &h, bf16x8v &l) {
            h = round_bf16(x0, x1);
            const f32x4v r0 = {x0[0] - (float)h[0], x0[1] - (float)h[1], x0[2] - (float)h[2], x0[3] - (float)h[3]};
            const f32x4v r1 = {x1[0] - (float)h[4], x1[1] - (float)h[5], x1[2] - (float)h[6], x1[3] - (float)h[7]};
            l = round_bf16(r0, r1);
        };
#pragma unroll
        for (int t = 0; t < 2; ++t) split(lo.a[t], hi.a[t], ah[t], al[t]);
#pragma unroll
        for (int c = 0; c < 2; ++c) split(lo.b[c], hi.b[c], bh[c], bl[c]);
        if (q + 1 < LF / 32) {
            ops_load<BKC>(lo, Xs, x0, Ls, 2 * q + 2, wave, lane);
            ops_load<BKC>(hi, Xs, x0, Ls, 2 * q + 3, wave, lane);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[t], bh[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bl[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bh[c], acc[t][c], 0, 0, 0);
            }
    }
}

/* BF16: the off-diagonal products X_p L_qp^T are BF16 products (FP32
 * accumulation) and the inverse products split-BF16 (mma_panel_bf16x3). */
template <bool BF16>
__global__ __launch_bounds__(256) void trsm_panel_kernel(PanelArgs p)
{
    __shared__ __attribute__((aligned(16))) float Xs[RM * XP];
    __shared__ __attribute__((aligned(16))) float Ls[LF * SP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * RM;
    const int rows = (int)min<int64_t>(RM, p.mb - r0);
    const int nq = (p.m + LF - 1) / LF;
    const int wm = nq * LF;   /* slab columns in use */

    /* steps in order: for each leaf q, the tiles L_qp (p < q), then inv_qq */
    const int nsteps = nq * (nq + 1) / 2;
    auto step_of = [&](int j, int &q, int &pp) {
        q = 0;
        while (j > q) { j -= q + 1; ++q; }
        pp = j;              /* pp == q: the inverse */
    };
    auto blk_of = [&](int q, int pp) {
        const int mq = min(LF, p.m - q * LF);
        if (pp == q) return Blk{p.inv + (int64_t)q * LF * LF, 0, LF, 1, mq, mq};
        const RfpView v = p.L.sub((int64_t)q * LF, (int64_t)pp * LF);
        return Blk{p.a, v.off, v.rs, v.cs, mq, LF};
    };

    /* one operand in flight: the next step's tile loads while this step's
     * products run */
    Staged st;
    bool kcn;
    {
        int q, pp;
        step_of(0, q, pp);
        const Blk b = blk_of(q, pp);
        kcn = b.cs == 1;
        st.load(b, tid);
    }

    /* slab: rows r0 .. r0 + rows, columns 0 .. m of X (zero beyond); all
     * loads issued before the LDS stores */
    constexpr int SL = panel_shape::SL;   /* 16-byte pieces per thread */
    const bool xc = p.X.cs == 1;
    {
        f32x4v v[SL];
#pragma unroll
        for (int e = 0; e < SL; ++e) {
            const int idx = tid + 256 * e;
            v[e] = f32x4v{0.f, 0.f, 0.f, 0.f};
            if (xc) {
                /* rows contiguous in memory: 16-byte loads along columns */
                const int r = idx / (MAXQ * LF / 4), c = (idx % (MAXQ * LF / 4)) * 4;
                if (r < rows && c < p.m) {
                    const float *src = p.a + p.X.off + (r0 + r) * p.X.rs + c;
                    if (c + 4 <= p.m) {
                        const f32x4u x = *(const f32x4u *)src;
                        v[e] = f32x4v{x[0], x[1], x[2], x[3]};
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q) v[e][q] = c + q < p.m ? src[q] : 0.f;
                    }
                }
            } else {
                /* columns contiguous: 16-byte loads along rows */
                const int r = (idx % (RM / 4)) * 4, c = idx / (RM / 4);
                if (c < p.m && r < rows) {
                    const float *src = p.a + p.X.off + (r0 + r) + (int64_t)c * p.X.cs;
                    if (r + 4 <= rows) {
                        const f32x4u x = *(const f32x4u *)src;
                        v[e] = f32x4v{x[0], x[1], x[2], x[3]};
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q) v[e][q] = r + q < rows ? src[q] : 0.f;
                    }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < SL; ++e) {
            const int idx = tid + 256 * e;
            if (xc) {
                const int r = idx / (MAXQ * LF / 4), c = (idx % (MAXQ * LF / 4)) * 4;
                *(f32x4v *)(Xs + r * XP + c) = v[e];
            } else {
                const int r = (idx % (RM / 4)) * 4, c = idx / (RM / 4);
#pragma unroll
                for (int q = 0; q < 4; ++q) Xs[(r + q) * XP + c] = v[e][q];
            }
        }
    }

    f32x4v acc[2][2];
    auto zero = [&]() {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[t][c][r] = 0.f;
    };
    zero();
    for (int j = 0; j < nsteps; ++j) {
        int q, pp;
        step_of(j, q, pp);
        const bool kc = kcn;
        lds_barrier();     /* the previous products are done with Ls (and the slab writes are visible) */
        st.store(Ls, SP, tid);
        if (pp == q) {
            /* R_q = X_q - acc in the slab */
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = t * 16 + 4 * (lane >> 4) + r;
                        const int col = q * LF + col_tile(wave, c) * 16 + (lane & 15);
                        Xs[row * XP + col] -= acc[t][c][r];
                    }
            zero();
        }
        lds_barrier();
        if (j + 1 < nsteps) {
            int q2, p2;
            step_of(j + 1, q2, p2);
            const Blk b = blk_of(q2, p2);
            kcn = b.cs == 1;
            st.load(b, tid);
        }
        if (pp < q) {
            if (BF16) {
                if (kc) mma_panel_bf16<true>(acc, Xs, pp * LF, Ls, wave, lane);
                else    mma_panel_bf16<false>(acc, Xs, pp * LF, Ls, wave, lane);
            } else {
                if (kc) mma_panel<true>(acc, Xs, pp * LF, Ls, wave, lane);
                else    mma_panel<false>(acc, Xs, pp * LF, Ls, wave, lane);
            }
        } else {
            /* X_q = R_q inv(L_qq)^T; the slab's leaf q is rewritten once every wave has read it */
            if (BF16) mma_panel_bf16x3<true>(acc, Xs, q * LF, Ls, wave, lane);
            else      mma_panel<true>(acc, Xs, q * LF, Ls, wave, lane);
            lds_barrier();
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = t * 16 + 4 * (lane >> 4) + r;
                        const int col = q * LF + col_tile(wave, c) * 16 + (lane & 15);
                        Xs[row * XP + col] = acc[t][c][r];
                    }
            zero();
        }
    }
    lds_barrier();

    /* slab -> X (and the BF16 shadow): LDS reads first, then the stores */
    {
        f32x4v v[SL];
#pragma unroll
        for (int e = 0; e < SL; ++e) {
            const int idx = tid + 256 * e;
            if (xc) {
                const int r = idx / (MAXQ * LF / 4), c = (idx % (MAXQ * LF / 4)) * 4;
                v[e] = *(const f32x4v *)(Xs + r * XP + c);
            } else {
                const int r = (idx % (RM / 4)) * 4, c = idx / (RM / 4);
#pragma unroll
                for (int q = 0; q < 4; ++q) v[e][q] = Xs[(r + q) * XP + c];
            }
        }
#pragma unroll
        for (int e = 0; e < SL; ++e) {
            const int idx = tid + 256 * e;
            if (xc) {
                const int r = idx / (MAXQ * LF / 4), c = (idx % (MAXQ * LF / 4)) * 4;
                if (r >= rows || c >= p.m) continue;
                float *dst = p.a + p.X.off + (r0 + r) * p.X.rs + c;
                if (c + 4 <= p.m) {
                    *(f32x4u *)dst = f32x4u{v[e][0], v[e][1], v[e][2], v[e][3]};
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (c + q < p.m) dst[q] = v[e][q];
                }
            } else {
                const int r = (idx % (RM / 4)) * 4, c = idx / (RM / 4);
                if (c >= p.m || r >= rows) continue;
                float *dst = p.a + p.X.off + (r0 + r) + (int64_t)c * p.X.cs;
                if (r + 4 <= rows) {
                    *(f32x4u *)dst = f32x4u{v[e][0], v[e][1], v[e][2], v[e][3]};
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (r + q < rows) dst[q] = v[e][q];
                }
            }
        }
    }
    if (p.sh_hi) {
        if (p.S.cs == 1 && p.S.rs % 8 == 0 && p.S.off % 8 == 0) {
            /* shadow rows k-contiguous and 16-byte aligned: eight columns per
             * thread and store (two 16-byte LDS reads), a scalar tail past m */
            constexpr int CH = MAXQ * LF / 8;   /* 8-column chunks per slab row */
#pragma unroll
            for (int e = 0; e < RM * CH / 256; ++e) {
                const int idx = tid + 256 * e, r = idx / CH, c = (idx % CH) * 8;
                if (r >= rows || c >= p.m) continue;
                const f32x4v x0 = *(const f32x4v *)(Xs + r * XP + c), x1 = *(const f32x4v *)(Xs + r * XP + c + 4);
                const float x[8] = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
                const int64_t so = p.S.off + (r0 + r) * p.S.rs + c;
                if (c + 8 <= p.m) {
                    bf16x8v hv, lv;
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        hv[q] = (__bf16)x[q];   /* nearest even */
                        lv[q] = (__bf16)(x[q] - (float)hv[q]);
                    }
                    *reinterpret_cast<bf16x8v *>(p.sh_hi + so) = hv;
                    if (p.sh_lo) *reinterpret_cast<bf16x8v *>(p.sh_lo + so) = lv;
                } else {
                    for (int q = 0; q < p.m - c; ++q) {
                        const unsigned short hb = bf16_rne(x[q]);
                        p.sh_hi[so + q] = hb;
                        if (p.sh_lo) p.sh_lo[so + q] = bf16_rne(x[q] - __uint_as_float((unsigned)hb << 16));
                    }
                }
            }
        } else {
            for (int idx = tid; idx < RM * wm; idx += 256) {
                const int r = idx / wm, c = idx % wm;
                if (r >= rows || c >= p.m) continue;
                const float v = Xs[r * XP + c];
                const int64_t so = p.S.off + (r0 + r) * p.S.rs + (int64_t)c * p.S.cs;
                const unsigned short hb = bf16_rne(v);
                p.sh_hi[so] = hb;
                if (p.sh_lo) p.sh_lo[so] = bf16_rne(v - __uint_as_float((unsigned)hb << 16));
            }
        }
    }
}

}  // namespace

int trsm_panel_max() { return MAXQ * LF; }

hipError_t launch_trsm_panel_f32(float *a, RfpView X, int64_t mb, RfpView L, int64_t m, const float *inv,
                                 unsigned short *sh_hi, unsigned short *sh_lo, RfpView S, bool bf16, hipStream_t s)
{
    if (mb <= 0 || m <= 0) return hipSuccess;
    if (m > MAXQ * LF || mb > (int64_t)INT32_MAX) return hipErrorInvalidValue;
    PanelArgs pa;
    pa.a = a;
    pa.X = X;
    pa.L = L;
    pa.inv = inv;
    pa.mb = (int)mb;
    pa.m = (int)m;
    pa.sh_hi = sh_hi;
    pa.sh_lo = sh_lo;
    pa.S = S;
    const unsigned grid = (unsigned)((mb + RM - 1) / RM);
    if (bf16) hipLaunchKernelGGL(trsm_panel_kernel<true>, dim3(grid), dim3(256), 0, s, pa);
    else      hipLaunchKernelGGL(trsm_panel_kernel<false>, dim3(grid), dim3(256), 0, s, pa);
    return hipGetLastError();
}
