/*
 * pfcon.hip — the vector kernels of the reciprocal condition estimate from an
 * RFP Cholesky factor (reference src/curfp_spfcon.cu:33-79): filling the
 * iterate (1/n, a unit vector e_j, or LAPACK ?lacn2's alternating vector
 * (-1)^i (1 + i/(n-1))), replacing it by its sign pattern, and one reduction
 * that yields everything the estimator's next decision needs in a single
 * 48-byte record — sum |x_i|, sum isgn_i x_i, max |x_i| and its first index,
 * and |x_j| for the previous index j — so each iteration costs one solve,
 * one reduction and one small read-back.  The reference takes three cuBLAS
 * reductions and two element copies for the same decisions.
 *
 * Both sums use the same per-thread order and the same tree, so
 * sum isgn_i x_i == sum |x_i| holds bit for bit exactly when every sign
 * agrees, the test the estimator makes.
 */
#include "rfp_internal.h"

namespace {

constexpr int RT = 1024;   /* threads of the reduction workgroup */

template <typename T>
__global__ __launch_bounds__(256) void pfcon_fill_kernel(T *x, int64_t n, int mode, int64_t j)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v;
    if (mode == 0)      v = T(1) / T(n);
    else if (mode == 1) v = (i == j) ? T(1) : T(0);
    else                v = ((i & 1) ? T(-1) : T(1)) * (n > 1 ? T(1) + T(i) / T(n - 1) : T(1));
    x[i] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void pfcon_sign_kernel(T *x, T *isgn, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T s = x[i] >= T(0) ? T(1) : T(-1);
    x[i] = s;
    isgn[i] = s;
}

template <typename T>
__global__ __launch_bounds__(RT) void pfcon_stats_kernel(const T *x, const T *isgn, int64_t n, int64_t j,
                                                         PfconStats *out)
{
    __shared__ double s_sum[RT / 64], s_dot[RT / 64], s_max[RT / 64];
    __shared__ int64_t s_idx[RT / 64];
    const int tid = threadIdx.x;
    double sum = 0.0, dot = 0.0, mx = -1.0;
    int64_t mi = 0;
    for (int64_t i = tid; i < n; i += RT) {
        const double v = (double)x[i];
        const double a = fabs(v);
        sum += a;
        if (isgn) dot += (double)isgn[i] * v;
        if (a > mx) { mx = a; mi = i; }
    }
    /* wave level: the larger |x|, the smaller index on a tie */
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        sum += __shfl_xor(sum, o);
        dot += __shfl_xor(dot, o);
        const double om = __shfl_xor(mx, o);
        const int64_t oi = __shfl_xor(mi, o);
        if (om > mx || (om == mx && oi < mi)) { mx = om; mi = oi; }
    }
    const int w = tid >> 6;
    if ((tid & 63) == 0) { s_sum[w] = sum; s_dot[w] = dot; s_max[w] = mx; s_idx[w] = mi; }
    __syncthreads();
    if (tid == 0) {
        sum = s_sum[0]; dot = s_dot[0]; mx = s_max[0]; mi = s_idx[0];
        for (int k = 1; k < RT / 64; ++k) {
            sum += s_sum[k];
            dot += s_dot[k];
            if (s_max[k] > mx || (s_max[k] == mx && s_idx[k] < mi)) { mx = s_max[k]; mi = s_idx[k]; }
        }
        PfconStats r;
        r.asum = sum;
        r.dot = dot;
        r.xmax = mx;
        r.xj = (j >= 0 && j < n) ? fabs((double)x[j]) : 0.0;
        r.imax = mi;
        r.pad = 0;
        *out = r;
    }
}

template <typename T>
hipError_t fill_t(T *x, int64_t n, int mode, int64_t j, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((pfcon_fill_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, mode, j);
    return hipGetLastError();
}

template <typename T>
hipError_t sign_t(T *x, T *isgn, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((pfcon_sign_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, isgn, n);
    return hipGetLastError();
}

template <typename T>
hipError_t stats_t(const T *x, const T *isgn, int64_t n, int64_t j, PfconStats *out, hipStream_t s)
{
    hipLaunchKernelGGL((pfcon_stats_kernel<T>), dim3(1), dim3(RT), 0, s, x, isgn, n, j, out);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_pfcon_fill_f32(float *x, int64_t n, int mode, int64_t j, hipStream_t s) { return fill_t(x, n, mode, j, s); }
hipError_t launch_pfcon_fill_f64(double *x, int64_t n, int mode, int64_t j, hipStream_t s) { return fill_t(x, n, mode, j, s); }
hipError_t launch_pfcon_sign_f32(float *x, float *isgn, int64_t n, hipStream_t s) { return sign_t(x, isgn, n, s); }
hipError_t launch_pfcon_sign_f64(double *x, double *isgn, int64_t n, hipStream_t s) { return sign_t(x, isgn, n, s); }
hipError_t launch_pfcon_stats_f32(const float *x, const float *isgn, int64_t n, int64_t j, PfconStats *out,
                                  hipStream_t s)
{
    return stats_t(x, isgn, n, j, out, s);
}
hipError_t launch_pfcon_stats_f64(const double *x, const double *isgn, int64_t n, int64_t j, PfconStats *out,
                                  hipStream_t s)
{
    return stats_t(x, isgn, n, j, out, s);
}
