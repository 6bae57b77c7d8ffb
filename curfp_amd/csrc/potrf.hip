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

template <typename T> struct LeafSize;
template <> struct LeafSize<float>  { static constexpr int value = 128; };
template <> struct LeafSize<double> { static constexpr int value = 64; };

int rfp_leaf(bool f64) { return f64 ? LeafSize<double>::value : LeafSize<float>::value; }

/* Load the lower triangle of an m x m view into s (pitch P), zero above. */
template <typename T, int NB, int P>
__device__ __forceinline__ void load_lower(T *s, const T *base, RfpView v, int m, int tid)
{
    const int total = NB * NB;
    if (v.rs == 1) {
        for (int idx = tid; idx < total; idx += 256) {
            const int i = idx % NB, j = idx / NB;
            T x = 0;
            if (i < m && j < m && i >= j) x = base[v.off + i + (int64_t)j * v.cs];
            s[i * P + j] = x;
        }
    } else {
        for (int idx = tid; idx < total; idx += 256) {
            const int j = idx % NB, i = idx / NB;
            T x = 0;
            if (i < m && j < m && i >= j) x = base[v.off + (int64_t)i * v.rs + j * v.cs];
            s[i * P + j] = x;
        }
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
__global__ __launch_bounds__(256) void trtri_tiles_kernel(const T *base, RfpView v, int64_t m, T *inv)
{
    constexpr int P = NB + 1;
    constexpr int RB = NB / 16;
    __shared__ T s[NB * P];
    __shared__ T colbuf[2][NB];
    __shared__ T rowbuf[2][NB];
    const int tid = threadIdx.x;
    const int64_t t = blockIdx.x;
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

template <typename T>
static hipError_t potrf_tile(T *base, RfpView v, int m, int *info, int info_base, T *inv, hipStream_t s)
{
    constexpr int NB = LeafSize<T>::value;
    if (m <= 0) return hipSuccess;
    if (m > NB) return hipErrorInvalidValue;
    hipLaunchKernelGGL((potrf_tile_kernel<T, NB>), dim3(1), dim3(256), 0, s, base, v, m, info, info_base, inv);
    return hipGetLastError();
}

template <typename T>
static hipError_t trtri_tiles(const T *base, RfpView v, int64_t m, T *inv, hipStream_t s)
{
    constexpr int NB = LeafSize<T>::value;
    if (m <= 0) return hipSuccess;
    const int64_t nt = (m + NB - 1) / NB;
    hipLaunchKernelGGL((trtri_tiles_kernel<T, NB>), dim3((unsigned)nt), dim3(256), 0, s, base, v, m, inv);
    return hipGetLastError();
}

hipError_t launch_potrf_tile_f32(float *base, RfpView v, int m, int *info, int info_base, float *inv, hipStream_t s)
{ return potrf_tile<float>(base, v, m, info, info_base, inv, s); }
hipError_t launch_potrf_tile_f64(double *base, RfpView v, int m, int *info, int info_base, double *inv, hipStream_t s)
{ return potrf_tile<double>(base, v, m, info, info_base, inv, s); }
hipError_t launch_trtri_tiles_f32(const float *base, RfpView v, int64_t m, float *inv, hipStream_t s)
{ return trtri_tiles<float>(base, v, m, inv, s); }
hipError_t launch_trtri_tiles_f64(const double *base, RfpView v, int64_t m, double *inv, hipStream_t s)
{ return trtri_tiles<double>(base, v, m, inv, s); }
