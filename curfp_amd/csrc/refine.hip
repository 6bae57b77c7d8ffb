/*
 * refine.hip — small kernels of the mixed-precision iterative refinement
 * (curfpSpftrsRefine): FP64 <-> FP32 conversions of the n x nrhs residual and
 * correction, the FP64 correction accumulation X += D, and FP64 sum-of-squares
 * reductions for the stopping test.  Everything here is a pure bandwidth
 * kernel over n x nrhs (or k x k) elements.
 */
#include "rfp_internal.h"

namespace {

/* The element-wise kernels below run on a 2-D grid: x over the n rows, y
 * over the columns (strided when nrhs exceeds the grid), so no thread
 * divides a flat index by n. */
__global__ void cast_f64_f32_kernel(int64_t n, int64_t nrhs, const double *src, int64_t lds, float *dst,
                                    int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int64_t c = blockIdx.y; c < nrhs; c += gridDim.y) dst[i + c * ldd] = (float)src[i + c * lds];
}

/* X = D (accumulate = false) or X += D, D in FP32, X in FP64 */
__global__ void acc_f32_f64_kernel(int64_t n, int64_t nrhs, const float *D, int64_t ldd, double *X, int64_t ldx,
                                   int accumulate)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int64_t c = blockIdx.y; c < nrhs; c += gridDim.y) {
        const double d = (double)D[i + c * ldd];
        double *x = X + i + c * ldx;
        *x = accumulate ? *x + d : d;
    }
}

/* Y = X, both n x nrhs column-major FP64 */
__global__ void copy_f64_kernel(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *Y, int64_t ldy)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int64_t c = blockIdx.y; c < nrhs; c += gridDim.y) Y[i + c * ldy] = X[i + c * ldx];
}

/* out[slot] += sum of squares of an n x nrhs column-major FP64 matrix */
__global__ void sumsq_kernel(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t c = blockIdx.y; c < nrhs; c += gridDim.y)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
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

/* ---- flexible GMRES on n x nrhs blocks (one Krylov space per column) ----
 * Vector j of a basis is the column-major n x nrhs FP64 block at
 * base + j * n * nrhs; every column c of every kernel below is independent. */

/* Partial dot products of one GM_CHUNK-row piece of column c (blockIdx.y)
 * with each of nv basis vectors: part[(chunk * nv + i) * nrhs + c] =
 * sum_r V_i(r, c) w(r, c).  w's piece is staged in LDS once. */
constexpr int GM_CHUNK = 2048;
__global__ __launch_bounds__(256) void gm_dots_kernel(int64_t n, int64_t nrhs, const double *V, int nv,
                                                      const double *w, double *part)
{
    __shared__ double ws[GM_CHUNK];
    __shared__ double red[4];
    const int64_t c = blockIdx.y, r0 = (int64_t)blockIdx.x * GM_CHUNK;
    const int rows = (int)min<int64_t>(GM_CHUNK, n - r0);
    const int tid = threadIdx.x;
    const double *wc = w + c * n + r0;
    for (int r = tid; r < rows; r += 256) ws[r] = wc[r];
    __syncthreads();
    for (int i = 0; i < nv; ++i) {
        const double *vc = V + (int64_t)i * n * nrhs + c * n + r0;
        double sum = 0.0;
        for (int r = tid; r < rows; r += 256) sum += vc[r] * ws[r];
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if ((tid & 63) == 0) red[tid >> 6] = sum;
        __syncthreads();
        if (tid == 0) part[((int64_t)blockIdx.x * nv + i) * nrhs + c] = (red[0] + red[1]) + (red[2] + red[3]);
        __syncthreads();
    }
}

/* out[t] = sum over chunks of part[chunk * cnt + t], in chunk order */
__global__ void gm_reduce_kernel(int nchunks, int64_t cnt, const double *part, double *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    double s = 0.0;
    for (int ch = 0; ch < nchunks; ++ch) s += part[(int64_t)ch * cnt + t];
    out[t] = s;
}

/* w(r, c) -= sum_{i < nv} h[i * nrhs + c] V_i(r, c) */
__global__ void gm_update_kernel(int64_t n, int64_t nrhs, const double *V, int nv, const double *h, double *w)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t c = idx / n;
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s += h[(int64_t)i * nrhs + c] * V[(int64_t)i * n * nrhs + idx];
    w[idx] -= s;
}

/* v(r, c) = w(r, c) / sqrt(nrm2[c]) (zero for a zero column) */
__global__ void gm_normalize_kernel(int64_t n, int64_t nrhs, const double *w, const double *nrm2, double *v)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const double q = nrm2[idx / n];
    v[idx] = q > 0.0 ? w[idx] / sqrt(q) : 0.0;
}

/* X(r, c) += sum_{j < nz} y[j * nrhs + c] Z_j(r, c) */
__global__ void gm_combine_kernel(int64_t n, int64_t nrhs, const double *Z, int nz, const double *y, double *X,
                                  int64_t ldx)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t r = idx % n, c = idx / n;
    double s = 0.0;
    for (int j = 0; j < nz; ++j) s += y[(int64_t)j * nrhs + c] * Z[(int64_t)j * n * nrhs + idx];
    X[r + c * ldx] += s;
}

__global__ void zero_f64_kernel(double *p, int n)
{
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) p[i] = 0.0;
}

/* 2-D grid for an n x nrhs element-wise pass: rows over x, columns over y */
inline dim3 grid_2d(int64_t n, int64_t nrhs, int64_t xcap = 1LL << 30)
{
    int64_t bx = (n + 255) / 256;
    if (bx > xcap) bx = xcap;
    if (bx < 1) bx = 1;
    const int64_t by = nrhs < 1 ? 1 : (nrhs > 65535 ? 65535 : nrhs);
    return dim3((unsigned)bx, (unsigned)by);
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
    hipLaunchKernelGGL(cast_f64_f32_kernel, grid_2d(n, nrhs), dim3(256), 0, s, n, nrhs, src,
                       lds, dst, ldd);
    return hipGetLastError();
}

hipError_t launch_acc_f32_f64(int64_t n, int64_t nrhs, const float *D, int64_t ldd, double *X, int64_t ldx,
                              bool accumulate, hipStream_t s)
{
    hipLaunchKernelGGL(acc_f32_f64_kernel, grid_2d(n, nrhs), dim3(256), 0, s, n, nrhs, D, ldd,
                       X, ldx, accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_copy_f64(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *Y, int64_t ldy,
                           hipStream_t s)
{
    hipLaunchKernelGGL(copy_f64_kernel, grid_2d(n, nrhs), dim3(256), 0, s, n, nrhs, X, ldx, Y,
                       ldy);
    return hipGetLastError();
}

hipError_t launch_sumsq(int64_t n, int64_t nrhs, const double *X, int64_t ldx, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(sumsq_kernel, grid_2d(n, nrhs < 64 ? nrhs : 64, 32), dim3(256), 0, s, n, nrhs, X, ldx, out);
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

int64_t gm_chunks(int64_t n) { return (n + GM_CHUNK - 1) / GM_CHUNK; }

hipError_t launch_gm_dots(int64_t n, int64_t nrhs, const double *V, int nv, const double *w, double *part,
                          double *out, hipStream_t s)
{
    const int64_t nch = gm_chunks(n);
    if (nv <= 0 || nrhs > 65535 || nch > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gm_dots_kernel, dim3((unsigned)nch, (unsigned)nrhs), dim3(256), 0, s, n, nrhs, V, nv, w, part);
    const int64_t cnt = (int64_t)nv * nrhs;
    hipLaunchKernelGGL(gm_reduce_kernel, dim3(blocks_for(cnt)), dim3(256), 0, s, (int)nch, cnt, part, out);
    return hipGetLastError();
}

hipError_t launch_gm_update(int64_t n, int64_t nrhs, const double *V, int nv, const double *h, double *w,
                            hipStream_t s)
{
    hipLaunchKernelGGL(gm_update_kernel, dim3(blocks_for(n * nrhs, 1LL << 30)), dim3(256), 0, s, n, nrhs, V, nv, h, w);
    return hipGetLastError();
}

hipError_t launch_gm_normalize(int64_t n, int64_t nrhs, const double *w, const double *nrm2, double *v, hipStream_t s)
{
    hipLaunchKernelGGL(gm_normalize_kernel, dim3(blocks_for(n * nrhs, 1LL << 30)), dim3(256), 0, s, n, nrhs, w, nrm2,
                       v);
    return hipGetLastError();
}

hipError_t launch_gm_combine(int64_t n, int64_t nrhs, const double *Z, int nz, const double *y, double *X, int64_t ldx,
                             hipStream_t s)
{
    hipLaunchKernelGGL(gm_combine_kernel, dim3(blocks_for(n * nrhs, 1LL << 30)), dim3(256), 0, s, n, nrhs, Z, nz, y, X,
                       ldx);
    return hipGetLastError();
}
