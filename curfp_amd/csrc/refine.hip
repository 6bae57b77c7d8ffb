/*
 * refine.hip — small kernels of the mixed-precision iterative refinement
 * (curfpSpftrsRefine): FP64 <-> FP32 conversions of the n x nrhs residual and
 * correction, the FP64 correction accumulation X += D, and FP64 sum-of-squares
 * reductions for the stopping test.  Everything here is a pure bandwidth
 * kernel over n x nrhs (or k x k) elements.
 */
#include "rfp_internal.h"

namespace {

__global__ void cast_f64_f32_kernel(int64_t n, int64_t nrhs, const double *src, int64_t lds, float *dst,
                                    int64_t ldd)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx % n, c = idx / n;
    dst[i + c * ldd] = (float)src[i + c * lds];
}

/* X = D (accumulate = false) or X += D, D in FP32, X in FP64 */
__global__ void acc_f32_f64_kernel(int64_t n, int64_t nrhs, const float *D, int64_t ldd, double *X, int64_t ldx,
                                   int accumulate)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx % n, c = idx / n;
    const double d = (double)D[i + c * ldd];
    double *x = X + i + c * ldx;
    *x = accumulate ? *x + d : d;
}

/* out[slot] += sum of squares of an n x nrhs column-major FP64 matrix */
__global__ void sumsq_kernel(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *out)
{
    __shared__ double red[256];
    const int64_t total = n * nrhs;
    double s = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % n, c = idx / n;
        const double v = X[i + c * ldx];
        s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

/* out[0] += sum G(i,j)^2, out[1] += trace G, for a k x k column-major G */
/* ||G||_F^2 and trace(G) of the symmetric k x k G from its lower triangle
 * (column-major; the strict upper part is not read) */
__global__ void gram_kernel(int64_t k, const double *G, double *out)
{
    __shared__ double r0[256], r1[256];
    double s = 0.0, t = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < k * k;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % k, j = idx / k;
        if (i < j) continue;
        const double v = G[idx];
        s += (i == j ? 1.0 : 2.0) * v * v;
        if (i == j) t += v;
    }
    r0[threadIdx.x] = s;
    r1[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(out, r0[0]);
        atomicAdd(out + 1, r1[0]);
    }
}

/* the same from an FP32 G (norms accumulated in FP64) */
__global__ void gram_f32_kernel(int64_t k, const float *G, double *out)
{
    __shared__ double r0[256], r1[256];
    double s = 0.0, t = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < k * k;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % k, j = idx / k;
        if (i < j) continue;
        const double v = (double)G[idx];
        s += (i == j ? 1.0 : 2.0) * v * v;
        if (i == j) t += v;
    }
    r0[threadIdx.x] = s;
    r1[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(out, r0[0]);
        atomicAdd(out + 1, r1[0]);
    }
}

__global__ void zero_f64_kernel(double *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

inline unsigned blocks_for(int64_t total, int64_t cap = 1 << 20)
{
    int64_t b = (total + 255) / 256;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace

hipError_t launch_cast_f64_f32(int64_t n, int64_t nrhs, const double *src, int64_t lds, float *dst, int64_t ldd,
                               hipStream_t s)
{
    hipLaunchKernelGGL(cast_f64_f32_kernel, dim3(blocks_for(n * nrhs, 1LL << 30)), dim3(256), 0, s, n, nrhs, src,
                       lds, dst, ldd);
    return hipGetLastError();
}

hipError_t launch_acc_f32_f64(int64_t n, int64_t nrhs, const float *D, int64_t ldd, double *X, int64_t ldx,
                              bool accumulate, hipStream_t s)
{
    hipLaunchKernelGGL(acc_f32_f64_kernel, dim3(blocks_for(n * nrhs, 1LL << 30)), dim3(256), 0, s, n, nrhs, D, ldd,
                       X, ldx, accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_sumsq(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(sumsq_kernel, dim3(blocks_for(n * nrhs, 2048)), dim3(256), 0, s, n, nrhs, X, ldx, out);
    return hipGetLastError();
}

hipError_t launch_gram_norms_f32(int64_t k, const float *G, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(gram_f32_kernel, dim3(blocks_for(k * k, 1024)), dim3(256), 0, s, k, G, out);
    return hipGetLastError();
}

hipError_t launch_gram_norms(int64_t k, const double *G, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(gram_kernel, dim3(blocks_for(k * k, 1024)), dim3(256), 0, s, k, G, out);
    return hipGetLastError();
}

hipError_t launch_zero_f64(double *p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}
