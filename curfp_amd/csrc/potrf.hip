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
 * In-place inverse of the m x m lower-triangular matrix in s (pitch P),
 * by rows: X(i, c) = -(sum_{c <= k < i} L(i, k) X(k, c)) / L(i, i).
 * Row i of L is consumed while row i of X is produced, so the new row goes
 * through `row` first.  Two threads share each column c (halves of the k
 * range) and combine through `part`.
 */
template <typename T, int NB, int P>
__device__ void invert_lower(T *s, T *row, T *part, int m, int tid)
{
    const int c = tid % NB;
    const int half = tid / NB;           /* 0 or 1 when NB == 128; NB == 64 uses 4 quarters */
    constexpr int SPLIT = 256 / NB;
    for (int i = 0; i < m; ++i) {
        T sum = 0;
        if (c < i) {
            const int len = i - c;
            const int chunk = (len + SPLIT - 1) / SPLIT;
            const int k0 = c + half * chunk;
            const int k1 = min(i, k0 + chunk);
            for (int k = k0; k < k1; ++k) sum += s[i * P + k] * s[k * P + c];
        }
        part[half * NB + c] = sum;
        __syncthreads();
        if (half == 0 && c <= i) {
            const T lii = s[i * P + i];
            T v;
            if (c == i) {
                v = T(1) / lii;
            } else {
                T tot = 0;
#pragma unroll
                for (int h = 0; h < SPLIT; ++h) tot += part[h * NB + c];
                v = -tot / lii;
            }
            row[c] = v;
        }
        __syncthreads();
        if (half == 0 && c <= i) s[i * P + c] = row[c];
        __syncthreads();
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

template <typename T>
__device__ __forceinline__ T rsqrt_exact(T x) { return T(1) / sqrt(x); }

template <typename T, int NB>
__global__ __launch_bounds__(256) void potrf_tile_kernel(T *base, RfpView v, int m, int *info,
                                                         int info_base, T *inv)
{
    constexpr int P = NB + 1;
    __shared__ T s[NB * P];
    __shared__ T row[NB];
    __shared__ T part[256];
    __shared__ int fail;
    const int tid = threadIdx.x;
    if (*info != INT_MAX) return;          /* an earlier block already failed */

    load_lower<T, NB, P>(s, base, v, m, tid);
    if (tid == 0) fail = -1;
    __syncthreads();

    /* right-looking column Cholesky; 16 x 16 thread grid on the update */
    const int tx = tid & 15, ty = tid >> 4;
    for (int j = 0; j < m; ++j) {
        const T d = s[j * P + j];
        if (!(d > T(0))) {                 /* uniform: every thread read the same d */
            if (tid == 0) fail = j;
            break;
        }
        const T ljj = sqrt(d);
        const T rinv = T(1) / ljj;
        __syncthreads();                   /* all threads have read d */
        if (tid == 0) s[j * P + j] = ljj;
        for (int i = j + 1 + tid; i < m; i += 256) s[i * P + j] *= rinv;
        __syncthreads();
        for (int i = j + 1 + ty; i < m; i += 16) {
            const T lij = s[i * P + j];
            for (int k = j + 1 + tx; k <= i; k += 16) s[i * P + k] -= lij * s[k * P + j];
        }
        __syncthreads();
    }
    __syncthreads();

    store_lower<T, NB, P>(s, base, v, m, tid);
    if (fail >= 0) {
        if (tid == 0) atomicMin(info, info_base + fail + 1);
        return;
    }
    if (inv) {
        __syncthreads();
        invert_lower<T, NB, P>(s, row, part, m, tid);
        write_inverse<T, NB, P>(s, inv, m, tid);
    }
}

template <typename T, int NB>
__global__ __launch_bounds__(256) void trtri_tiles_kernel(const T *base, RfpView v, int64_t m, T *inv)
{
    constexpr int P = NB + 1;
    __shared__ T s[NB * P];
    __shared__ T row[NB];
    __shared__ T part[256];
    const int tid = threadIdx.x;
    const int64_t t = blockIdx.x;
    const int64_t r0 = t * NB;
    const int mt = (int)((m - r0) < NB ? (m - r0) : NB);
    load_lower<T, NB, P>(s, base, v.sub(r0, r0), mt, tid);
    __syncthreads();
    invert_lower<T, NB, P>(s, row, part, mt, tid);
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
