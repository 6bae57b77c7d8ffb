/*
 * trsm_panel64.hip — FP64 twin of the panel TRSM at the bottom of dpftrf's
 * recursive triangular solve: one launch for X := X * L^{-T} with L a
 * lower-triangular diagonal block of at most four 64-row leaves (the
 * reference's TRSM step of the RFP Cholesky, src/curfp_dpftrf.cpp:140-244,
 * reached here through the recursion over the three RFP views).
 *
 * Each workgroup keeps 32 rows of X (a slab) in LDS and walks the leaves:
 *
 *     X_q := (X_q - sum_{p < q} X_p L_qp^T) * inv(L_qq)^T
 *
 * with the leaf inverses written by the FP64 leaf Cholesky (row-major 64 x 64,
 * zero above the diagonal).  The L tiles and inverses are staged through
 * registers into two LDS buffers, the next tile's loads in flight during the
 * current tile's products; products are v_mfma_f64_16x16x4_f64.  Of the 32 x 64
 * output of a leaf each wave owns two 16 x 16 tiles, row tile 0 with column
 * tile w and row tile 1 with column tile 3 - w, so the triangular inverse
 * products (column tile c needs k < 16 (c + 1)) give every wave 20 k steps.
 */
#include "rfp_internal.h"
#include "trsm_panel_shape.h"

typedef double f64x4 __attribute__((ext_vector_type(4)));

namespace {

using namespace panel64_shape;

struct Panel64Args {
    double *a;              /* RFP array holding X and L */
    RfpView X;              /* mb x m */
    RfpView L;              /* m x m lower */
    const double *inv;      /* inverse of leaf q at inv + q*LF*LF, row-major */
    int mb, m;
};

__device__ __forceinline__ f64x4 mma(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

/* Step j of the panel's walk -> leaf q and tile column pp (pp == q: the
 * inverse of leaf q); steps run q = 0, 1, ... and within a leaf pp = 0..q. */
__device__ __forceinline__ void step_of(int j, int &q, int &pp)
{
    q = 0;
    pp = j;
    while (pp > q) { pp -= q + 1; ++q; }
}

/* The operand tile of step j into registers: T(n, k) = L(64q + n, 64pp + k)
 * (pp < q) or inv_q(n, k) (pp == q), zero outside the panel. */
__device__ __forceinline__ void tile_load(const Panel64Args &p, int j, int tid, double (&v)[TE])
{
    int q, pp;
    step_of(j, q, pp);
    const int mq = min(LF, p.m - q * LF);
    if (pp == q) {
        const double *src = p.inv + (int64_t)q * LF * LF;
#pragma unroll
        for (int e = 0; e < TE; ++e) {
            int c, o;
            tile_elem(tid, e, c, o);               /* c = k, o = n */
            v[e] = (o < mq && c < mq) ? src[o * LF + c] : 0.0;
        }
    } else {
        const bool kc = p.L.cs == 1;
        const RfpView t = p.L.sub((int64_t)q * LF, (int64_t)pp * LF);
#pragma unroll
        for (int e = 0; e < TE; ++e) {
            int c, o;
            tile_elem(tid, e, c, o);
            const int n = kc ? o : c, k = kc ? c : o;
            v[e] = n < mq ? p.a[t.at(n, k)] : 0.0;
        }
    }
}

/* registers -> LDS as T(n, k) at n * TP + k (the inverse is row-major, so
 * k is its contiguous direction whatever L's layout) */
__device__ __forceinline__ void tile_put(const Panel64Args &p, int j, int tid, const double (&v)[TE], double *Ts)
{
    int q, pp;
    step_of(j, q, pp);
    const bool kc = pp == q || p.L.cs == 1;
#pragma unroll
    for (int e = 0; e < TE; ++e) {
        int c, o;
        tile_elem(tid, e, c, o);
        Ts[(kc ? o : c) * TP + (kc ? c : o)] = v[e];
    }
}

__global__ __launch_bounds__(THREADS) void trsm_panel64_kernel(Panel64Args p)
{
    __shared__ double Xs[RM * XP];
    __shared__ double Ts[2][LF * TP];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * RM;
    const int nq = (p.m + LF - 1) / LF, nsteps = nq * (nq + 1) / 2;
    const int ct[2] = {wave, 3 - wave};

    /* the slab: rows r0.., columns 0..m-1 (zero beyond mb and m).  All of a
     * thread's loads are issued before its LDS stores: an element outside
     * the panel is loaded from X(r0, 0) instead (inside the panel) and not
     * used, so no load waits in a branch of its own. */
    const bool xc = p.X.cs == 1;
    {
        double xv[SE];
#pragma unroll
        for (int e = 0; e < SE; ++e) {
            int r, c;
            slab_elem(tid, e, xc, r, c);
            const bool in = r0 + r < p.mb && c < p.m;
            xv[e] = p.a[in ? p.X.at(r0 + r, c) : p.X.at(r0, 0)];
        }
#pragma unroll
        for (int e = 0; e < SE; ++e) {
            int r, c;
            slab_elem(tid, e, xc, r, c);
            Xs[r * XP + c] = (r0 + r < p.mb && c < p.m) ? xv[e] : 0.0;
        }
    }
    double v[TE];
    tile_load(p, 0, tid, v);
    tile_put(p, 0, tid, v, Ts[0]);
    __syncthreads();

    f64x4 acc[2];
    for (int j = 0; j < nsteps; ++j) {
        int q, pp;
        step_of(j, q, pp);
        const double *T = Ts[j & 1];
        const bool more = j + 1 < nsteps;
        if (more) tile_load(p, j + 1, tid, v);

        if (pp < q) {
            if (pp == 0) {
                /* Y_q starts from X_q */
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        acc[t][r] = Xs[(16 * t + lk + 4 * r) * XP + q * LF + 16 * ct[t] + li];
            }
            /* Y_q -= X_pp L_{q,pp}^T */
            const int x0 = pp * LF;
#pragma unroll
            for (int k4 = 0; k4 < LF / 4; ++k4) {
                const int k = 4 * k4 + lk;
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    acc[t] = mma(-Xs[(16 * t + li) * XP + x0 + k], T[(16 * ct[t] + li) * TP + k], acc[t]);
            }
        } else {
            if (q > 0) {
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        Xs[(16 * t + lk + 4 * r) * XP + q * LF + 16 * ct[t] + li] = acc[t][r];
                __syncthreads();
            }
            /* X_q = Y_q inv(L_qq)^T; inv(n, k) = 0 for k > n */
            const int x0 = q * LF;
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k4 = 0; k4 < LF / 4; ++k4) {
                const int k = 4 * k4 + lk;
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    if (k4 < 4 * ct[t] + 4)
                        acc[t] = mma(Xs[(16 * t + li) * XP + x0 + k], T[(16 * ct[t] + li) * TP + k], acc[t]);
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Xs[(16 * t + lk + 4 * r) * XP + q * LF + 16 * ct[t] + li] = acc[t][r];
        }

        if (more) tile_put(p, j + 1, tid, v, Ts[(j + 1) & 1]);
        __syncthreads();
    }

    /* the slab back to X */
#pragma unroll 4
    for (int e = 0; e < SE; ++e) {
        int r, c;
        slab_elem(tid, e, xc, r, c);
        if (r0 + r < p.mb && c < p.m) p.a[p.X.at(r0 + r, c)] = Xs[r * XP + c];
    }
}

}  // namespace

int trsm_panel64_max() { return MAXQ * LF; }

hipError_t launch_trsm_panel_f64(double *a, RfpView X, int64_t mb, RfpView L, int64_t m, const double *inv,
                                 hipStream_t s)
{
    if (mb <= 0 || m <= 0) return hipSuccess;
    if (m > MAXQ * LF || mb > (int64_t)INT32_MAX) return hipErrorInvalidValue;
    /* both views stream along a unit stride */
    if ((X.cs != 1 && X.rs != 1) || (L.cs != 1 && L.rs != 1)) return hipErrorInvalidValue;
    Panel64Args pa;
    pa.a = a;
    pa.X = X;
    pa.L = L;
    pa.inv = inv;
    pa.mb = (int)mb;
    pa.m = (int)m;
    const unsigned grid = (unsigned)((mb + RM - 1) / RM);
    hipLaunchKernelGGL(trsm_panel64_kernel, dim3(grid), dim3(THREADS), 0, s, pa);
    return hipGetLastError();
}
