/*
 * lansf.hip — norms of a symmetric matrix stored in RFP: max |a_ij|, the
 * 1-norm (the largest absolute column sum of the full symmetric matrix, which
 * is also its infinity norm) and the Frobenius norm (reference
 * src/curfp_slansf.cu:273-435, src/curfp_dlansf.cu:189).  The reference reduces
 * with cuBLAS calls that take an int element count and returns
 * INVALID_VALUE once n(n+1)/2 > INT_MAX (n >= 65536); every index here is
 * 64-bit.
 *
 * max / Frobenius: one grid-stride pass over the packed array (16-byte loads,
 * FP64 sums), then a pass over the n diagonal entries: ||A||_F^2 =
 * 2 sum(arf^2) - sum(diag^2).
 *
 * 1-norm: the stored triangle L (lower, as the three views) gives the
 * absolute row sums of the full matrix as  s_i = sum_{j <= i} |L_ij| +
 * sum_{k > i} |L_ki|.  One workgroup per 128 x 128 tile (bi, bj), bi >= bj,
 * of the 128-row blocks of v11 then v22 (the block structure of the
 * substitution sweep) reads the tile once and writes its row sums to
 * S[bj][rows of bi] and its strict column sums to S[bi][cols of bj] (a
 * diagonal tile writes both into S[bi][block bi]) — the symmetric
 * matrix-vector kernel below with |L| and x = 1.  Every (slot, index) pair
 * of S is written by exactly one tile, so S needs no clearing and no
 * atomics; a second kernel sums the slots of each index in FP64 and takes
 * the maximum.
 */
#include "rfp_internal.h"
#include <algorithm>

namespace {

constexpr int NT = 128;       /* tile of the 1-norm */

struct LansfGeom {
    RfpView v11, v21, v22;
    int64_t n, n1, n2;
    int s1, nb;
};

__device__ __forceinline__ void lansf_block(const LansfGeom &g, int b, int64_t &grow, int64_t &lrow, int &rows)
{
    if (b < g.s1) {
        lrow = (int64_t)b * NT;
        grow = lrow;
        rows = (int)min<int64_t>(NT, g.n1 - lrow);
    } else {
        lrow = (int64_t)(b - g.s1) * NT;
        grow = g.n1 + lrow;
        rows = (int)min<int64_t>(NT, g.n2 - lrow);
    }
}

template <typename T> struct Vec16;
template <> struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); static constexpr int N = 4; };
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); static constexpr int N = 2; };

/* the order of non-negative IEEE values is the order of their bit patterns;
 * NaN's pattern is above +inf, so a NaN entry wins the maximum */
__device__ __forceinline__ unsigned long long key_of(float x) { return __float_as_uint(x); }
__device__ __forceinline__ unsigned long long key_of(double x) { return (unsigned long long)__double_as_longlong(x); }

template <typename T>
__device__ __forceinline__ T absmax(T m, T x)
{
    return (x > m || x != x) ? x : m;
}

/* block reduction of (max, sum) over 256 threads; thread 0 gets the result */
template <typename T>
__device__ __forceinline__ void block_reduce(T &m, double &s)
{
    __shared__ T sm[256];
    __shared__ double ss[256];
    const int tid = threadIdx.x;
    sm[tid] = m;
    ss[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            sm[tid] = absmax(sm[tid], sm[tid + w]);
            ss[tid] += ss[tid + w];
        }
        __syncthreads();
    }
    m = sm[0];
    s = ss[0];
}

/* acc[0] (bits) = max |a|, acc[1] (double) += sum a^2 over the nt entries */
template <typename T>
__global__ __launch_bounds__(256) void absmax_sumsq_kernel(const T *a, int64_t nt, unsigned long long *acc)
{
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    /* head up to a 16-byte boundary, 16-byte body, tail */
    const int64_t mis = (int64_t)(((uintptr_t)a / sizeof(T)) % VN);
    const int64_t head = min<int64_t>(nt, mis ? VN - mis : 0);
    const int64_t nv = (nt - head) / VN;
    T m = 0;
    double s = 0.0;
    if (tid < head) {
        const T x = fabs(a[tid]);
        m = absmax(m, x);
        s += (double)x * x;
    }
    const V *av = reinterpret_cast<const V *>(a + head);
    for (int64_t i = tid; i < nv; i += stride) {
        const V v = __builtin_nontemporal_load(av + i);   /* read once */
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            const T x = fabs(v[e]);
            m = absmax(m, x);
            s += (double)x * x;
        }
    }
    for (int64_t i = head + nv * VN + tid; i < nt; i += stride) {
        const T x = fabs(a[i]);
        m = absmax(m, x);
        s += (double)x * x;
    }
    block_reduce(m, s);
    if (threadIdx.x == 0) {
        atomicMax(acc, key_of(m));
        atomicAdd(reinterpret_cast<double *>(acc + 1), s);
    }
}

/* acc[2] (double) += sum of the squared diagonal */
template <typename T>
__global__ __launch_bounds__(256) void diag_sumsq_kernel(const T *a, LansfGeom g, unsigned long long *acc)
{
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += (int64_t)gridDim.x * blockDim.x) {
        const T x = a[i < g.n1 ? g.v11.at(i, i) : g.v22.at(i - g.n1, i - g.n1)];
        s += (double)x * x;
    }
    T m = 0;
    block_reduce(m, s);
    if (threadIdx.x == 0) atomicAdd(reinterpret_cast<double *>(acc + 2), s);
}

/* acc[3] (bits) = max_j sum_slot S[slot][j] */
template <typename T>
__global__ __launch_bounds__(256) void colsum_max_kernel(const T *S, LansfGeom g, unsigned long long *acc)
{
    T m = 0;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < g.n; j += (int64_t)gridDim.x * blockDim.x) {
        /* 16 independent loads in flight per thread */
        double part[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) part[u] = 0.0;
        int b = 0;
        for (; b + 16 <= g.nb; b += 16) {
            T x[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) x[u] = S[(int64_t)(b + u) * g.n + j];
#pragma unroll
            for (int u = 0; u < 16; ++u) part[u] += (double)x[u];
        }
        for (; b < g.nb; ++b) part[0] += (double)S[(int64_t)b * g.n + j];
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) s += part[u];
        m = absmax(m, (T)s);
    }
    double unused = 0.0;
    block_reduce(m, unused);
    if (threadIdx.x == 0) atomicMax(acc + 3, key_of(m));
}

/*
 * y := alpha A x + beta y with A symmetric in RFP: the tiles of the 1-norm's
 * traversal, each read once.  A tile L (rows of block bi, columns of block
 * bj, bi >= bj) contributes L x_bj to the rows of bi (slot bj) and, off the
 * diagonal, L^T x_bi to the columns of bj (slot bi); a diagonal tile writes
 * L x + strict(L)^T x into slot bi.  Every (slot, index) pair of S is
 * written by exactly one tile.  Tiles are taken strip by strip (all tiles of
 * column block 0 top to bottom, then column block 1, ...): consecutive
 * workgroups then read consecutive 128-row pieces of the same columns, which
 * measured 1.3x faster than the SYRK launches' banded order (DRAM page
 * locality; every byte is read once, so L2 reuse plays no part).  A thread's 16-byte piece runs along the
 * contiguous direction: its products summed within the piece give one
 * partial of the other direction (reduced over the RUN lanes that share the
 * piece's line), and the products themselves accumulate over the thread's
 * pieces for the VN lines of the contiguous direction (reduced over the
 * 256 / RUN thread groups through LDS).
 */
/* Tile t of the lower triangle in column-strip order: strip tj = 0, 1, ..,
 * and down each strip ti = tj, tj + 1, .. (consecutive workgroups read
 * consecutive 128-row pieces of the same columns). */
__device__ __forceinline__ void strip_tile_coords(int64_t t, int64_t T, int64_t &ti, int64_t &tj)
{
    const int64_t u = T * (T + 1) / 2 - 1 - t;     /* row-major index in the reversed triangle */
    int64_t r = (int64_t)((sqrt(8.0 * (double)u + 1.0) - 1.0) * 0.5);
    while (r > 0 && r * (r + 1) / 2 > u) --r;
    while ((r + 1) * (r + 2) / 2 <= u) ++r;
    const int64_t c = u - r * (r + 1) / 2;
    tj = T - 1 - r;
    ti = T - 1 - c;
}

template <typename T, bool ABS1>
__global__ __launch_bounds__(256) void symv_tiles_kernel(const T *a, LansfGeom g, const T *x, RfpView xv, T *S)
{
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    constexpr int PER = NT * NT / VN / 256;
    constexpr int RUN = NT / VN;         /* lanes along one line */
    constexpr int GROUPS = 256 / RUN;
    __shared__ T xr[NT], xc[NT], line[NT], red[GROUPS][NT];
    const int tid = threadIdx.x;
    int64_t bi, bj;
    strip_tile_coords((int64_t)blockIdx.x, g.nb, bi, bj);
    int64_t gr, lr, gc, lc;
    int mr, mc;
    lansf_block(g, (int)bi, gr, lr, mr);
    lansf_block(g, (int)bj, gc, lc, mc);
    const RfpView v = bi < g.s1 ? g.v11.sub(lr, lc) : (bj < g.s1 ? g.v21.sub(lr, lc) : g.v22.sub(lr, lc));
    const bool diag = bi == bj;
    const bool rc = v.rs == 1;            /* rows contiguous (column view) */
    const int a0 = (tid % RUN) * VN;      /* offset along the contiguous direction (fixed per thread) */
    T acc[VN];
#pragma unroll
    for (int k = 0; k < VN; ++k) acc[k] = T(0);
    const bool interior = !diag && mr == NT && mc == NT;
    /* every load of the tile issued before any arithmetic */
    T u_all[PER][VN];
    if (interior) {
        typedef T Vu __attribute__((ext_vector_type(VN), aligned(sizeof(T))));
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int b0 = tid / RUN + GROUPS * e;
            /* read once: non-temporal */
            const Vu w = __builtin_nontemporal_load((const Vu *)(a + (rc ? v.at(a0, b0) : v.at(b0, a0))));
#pragma unroll
            for (int k = 0; k < VN; ++k) u_all[e][k] = ABS1 ? fabs(w[k]) : w[k];
        }
    } else {
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int b0 = tid / RUN + GROUPS * e;
#pragma unroll
            for (int k = 0; k < VN; ++k) {
                const int r = rc ? a0 + k : b0, c = rc ? b0 : a0 + k;
                const T w = (r < mr && c < mc && (!diag || r >= c)) ? a[v.at(r, c)] : T(0);
                u_all[e][k] = ABS1 ? fabs(w) : w;
            }
        }
    }
    /* ABS1: the 1-norm's absolute row sums, x = 1 (not read) */
    if (tid < NT) xr[tid] = tid < mr ? (ABS1 ? T(1) : x[xv.off + (gr + tid) * xv.rs]) : T(0);
    else          xc[tid - NT] = tid - NT < mc ? (ABS1 ? T(1) : x[xv.off + (gc + tid - NT) * xv.rs]) : T(0);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int b0 = tid / RUN + GROUPS * e;     /* the line: a column (rc) or a row */
        const T *u = u_all[e];
        T part = 0;
        if (rc) {
            /* u[k] = L(a0 + k, b0): rows a0.. take L x; column b0 takes the
             * strict L^T x */
            const T xv_c = xc[b0];
#pragma unroll
            for (int k = 0; k < VN; ++k) {
                acc[k] += u[k] * xv_c;
                part += (diag && a0 + k == b0) ? T(0) : u[k] * xr[a0 + k];
            }
        } else {
            /* u[k] = L(b0, a0 + k): row b0 takes L x; columns a0.. the strict L^T x */
            const T xv_r = xr[b0];
#pragma unroll
            for (int k = 0; k < VN; ++k) {
                part += u[k] * xc[a0 + k];
                acc[k] += (diag && b0 == a0 + k) ? T(0) : u[k] * xv_r;
            }
        }
#pragma unroll
        for (int o = RUN / 2; o >= 1; o >>= 1) part += __shfl_xor(part, o);
        if (tid % RUN == 0) line[b0] = part;
    }
#pragma unroll
    for (int k = 0; k < VN; ++k) red[tid / RUN][a0 + k] = acc[k];
    __syncthreads();
    if (tid < NT) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < GROUPS; ++q) s += red[q][tid];
        /* rc: s is the row sum of row tid, line[] the column sums; else the
         * reverse */
        const T rsum = rc ? s : line[tid], csum = rc ? line[tid] : s;
        if (diag) {
            if (tid < mr) S[bi * g.n + gr + tid] = rsum + csum;
        } else {
            if (tid < mr) S[bj * g.n + gr + tid] = rsum;
            if (tid < mc) S[bi * g.n + gc + tid] = csum;
        }
    }
}

/* y(i) = alpha sum_slot S[slot][i] + beta y(i) (nslots = 0: alpha A x is 0;
 * beta = 0: y is not read) */
template <typename T>
__global__ __launch_bounds__(256) void symv_finish_kernel(const T *S, int64_t n, int nslots, T alpha, T beta, T *y,
                                                          RfpView yv)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        double part[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) part[u] = 0.0;
        int b = 0;
        for (; b + 8 <= nslots; b += 8) {
            T t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) t[u] = S[(int64_t)(b + u) * n + j];
#pragma unroll
            for (int u = 0; u < 8; ++u) part[u] += (double)t[u];
        }
        for (; b < nslots; ++b) part[0] += (double)S[(int64_t)b * n + j];
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) s += part[u];
        T *yp = y + yv.off + j * yv.rs;
        const T val = alpha * (T)s;
        *yp = beta != T(0) ? val + beta * *yp : val;
    }
}

LansfGeom lansf_geom(const RfpGeom &r);

template <typename T>
hipError_t symv(const RfpGeom &rg, const T *arf, const T *x, RfpView xv, T alpha, T beta, T *y, RfpView yv, void *ws,
                hipStream_t s)
{
    const LansfGeom g = lansf_geom(rg);
    T *S = (T *)ws;
    int nslots = 0;
    if (alpha != T(0)) {
        const int64_t ntiles = (int64_t)g.nb * (g.nb + 1) / 2;
        hipLaunchKernelGGL((symv_tiles_kernel<T, false>), dim3((unsigned)ntiles), dim3(256), 0, s, arf, g, x, xv, S);
        nslots = g.nb;
    }
    const unsigned blocks = (unsigned)std::min<int64_t>((g.n + 255) / 256, 2048);
    hipLaunchKernelGGL((symv_finish_kernel<T>), dim3(blocks), dim3(256), 0, s, (const T *)S, g.n, nslots, alpha, beta,
                       y, yv);
    return hipGetLastError();
}

__global__ void zero_u64_kernel(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

LansfGeom lansf_geom(const RfpGeom &r)
{
    LansfGeom g;
    g.v11 = r.v11; g.v21 = r.v21; g.v22 = r.v22;
    g.n = r.n; g.n1 = r.n1; g.n2 = r.n2;
    g.s1 = (int)((r.n1 + NT - 1) / NT);
    g.nb = g.s1 + (int)((r.n2 + NT - 1) / NT);
    return g;
}

template <typename T>
hipError_t lansf(const RfpGeom &rg, int norm, const T *arf, void *ws, hipStream_t s)
{
    const LansfGeom g = lansf_geom(rg);
    unsigned long long *acc = (unsigned long long *)ws;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, acc, 4);
    if (norm == 1) {
        T *S = (T *)((char *)ws + 256);
        const int64_t ntiles = (int64_t)g.nb * (g.nb + 1) / 2;
        hipLaunchKernelGGL((symv_tiles_kernel<T, true>), dim3((unsigned)ntiles), dim3(256), 0, s, arf, g, (const T *)nullptr,
                           RfpView{0, 1, 1}, S);
        const unsigned blocks = (unsigned)std::min<int64_t>((g.n + 255) / 256, 2048);
        hipLaunchKernelGGL((colsum_max_kernel<T>), dim3(blocks), dim3(256), 0, s, S, g, acc);
    } else {
        const int64_t nt = g.n * (g.n + 1) / 2;
        const unsigned blocks = (unsigned)std::min<int64_t>((nt / Vec16<T>::N + 255) / 256 + 1, 4096);
        hipLaunchKernelGGL((absmax_sumsq_kernel<T>), dim3(blocks), dim3(256), 0, s, arf, nt, acc);
        if (norm == 2) {
            const unsigned db = (unsigned)std::min<int64_t>((g.n + 255) / 256, 1024);
            hipLaunchKernelGGL((diag_sumsq_kernel<T>), dim3(db), dim3(256), 0, s, arf, g, acc);
        }
    }
    return hipGetLastError();
}

}  // namespace

size_t lansf_scratch_bytes(const RfpGeom &rg, int norm, size_t elem)
{
    const LansfGeom g = lansf_geom(rg);
    return 256 + (norm == 1 ? (size_t)g.nb * (size_t)g.n * elem : 0);
}

size_t symv_scratch_bytes(const RfpGeom &rg, size_t elem)
{
    const LansfGeom g = lansf_geom(rg);
    return (size_t)g.nb * (size_t)g.n * elem;
}

hipError_t launch_symv_f32(const RfpGeom &g, const float *arf, const float *x, RfpView xv, float alpha, float beta,
                           float *y, RfpView yv, void *ws, hipStream_t s)
{ return symv<float>(g, arf, x, xv, alpha, beta, y, yv, ws, s); }
hipError_t launch_symv_f64(const RfpGeom &g, const double *arf, const double *x, RfpView xv, double alpha,
                           double beta, double *y, RfpView yv, void *ws, hipStream_t s)
{ return symv<double>(g, arf, x, xv, alpha, beta, y, yv, ws, s); }

hipError_t launch_lansf_f32(const RfpGeom &g, int norm, const float *arf, void *ws, hipStream_t s)
{ return lansf<float>(g, norm, arf, ws, s); }
hipError_t launch_lansf_f64(const RfpGeom &g, int norm, const double *arf, void *ws, hipStream_t s)
{ return lansf<double>(g, norm, arf, ws, s); }
