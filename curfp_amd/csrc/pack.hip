/*
 * pack.hip — bandwidth kernels: RFP <-> dense triangle (strttf / stfttr and
 * the d* twins), add_to_diagonal, and small fills.
 *
 * Reference: src/curfp_strttf.cu:46-462 and src/curfp_stfttr.cu:24-175 launch
 * one thread per dense (row, col) and derive the packed index per case; the
 * transposed blocks then write (or read) the packed array with a large
 * stride.  Here each of the three RFP blocks (v11 lower triangle, v21
 * rectangle, v22 lower triangle) is copied by 64 x 64 tiles staged through
 * LDS: the source is read along its contiguous direction, the tile is
 * transposed in LDS when needed, and the destination is written along its
 * contiguous direction, so both sides stream coalesced — as 16-byte accesses
 * when a view's offset and leading dimension keep them aligned.  One launch
 * covers all three blocks; all offsets are 64-bit.
 */
#include "rfp_internal.h"

namespace {

constexpr int PT = 64;   /* tile edge */

struct CopySeg {
    RfpView src, dst;     /* element (i, j) of the block in each array */
    int64_t rows, cols;   /* block shape */
    int64_t tiles_r, tiles_c, first;  /* tile grid and first workgroup id */
    int tri;              /* lower-triangular block */
    int vsrc, vdst;       /* 16-byte accesses keep alignment on that side */
};

struct CopyParams {
    CopySeg seg[3];
    int nseg;
};

__device__ __forceinline__ void seg_tile(const CopySeg &s, int64_t t, int64_t &ti, int64_t &tj)
{
    if (s.tri) {
        int64_t r = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (r * (r + 1) / 2 > t) --r;
        while ((r + 1) * (r + 2) / 2 <= t) ++r;
        ti = r;
        tj = t - r * (r + 1) / 2;
    } else {
        ti = t % s.tiles_r;
        tj = t / s.tiles_r;
    }
}

/* 16-byte vector of T that only assumes element alignment of its type */
template <typename T> struct Vec16;
template <> struct Vec16<float>  { typedef float  type __attribute__((ext_vector_type(4))); };
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };

/* Thread tid's pieces of a PT x PT tile along a view's contiguous direction:
 * piece e covers VN consecutive elements starting at (li, lj) of the tile
 * (rows li.. when the direction is i, columns lj.. otherwise).  VN = 1 gives
 * the scalar layout (64 fast lanes x 4 slow rows). */
template <int VN>
struct TilePiece {
    static constexpr int PER_ROW = PT / VN;            /* threads along the contiguous direction */
    static constexpr int ROWS = 256 / PER_ROW;         /* slow rows per pass */
    static constexpr int PASSES = PT / ROWS;
    __device__ static void at(int tid, int e, bool along_i, int &li, int &lj)
    {
        const int c = (tid % PER_ROW) * VN, q = tid / PER_ROW + ROWS * e;
        li = along_i ? c : q;
        lj = along_i ? q : c;
    }
};

template <typename T>
__global__ __launch_bounds__(256) void rfp_copy_kernel(const T *src, T *dst, CopyParams p)
{
    __shared__ T tile[PT][PT + 1];
    constexpr int VN = 16 / sizeof(T);
    typedef typename Vec16<T>::type V;
    const int64_t bid = blockIdx.x;
    int si = 0;
    while (si + 1 < p.nseg && bid >= p.seg[si + 1].first) ++si;
    const CopySeg &s = p.seg[si];
    int64_t ti, tj;
    seg_tile(s, bid - s.first, ti, tj);
    const int64_t i0 = ti * PT, j0 = tj * PT;
    const int tid = threadIdx.x;
    /* element (gi, gj) of the segment lies in it (the lower part for a triangle) */
    auto inside = [&](int64_t gi, int64_t gj) { return gi < s.rows && gj < s.cols && (!s.tri || gi >= gj); };

    /* read along the source's contiguous direction; 16-byte loads when the
     * view's offset and leading dimension keep every piece aligned (whole
     * pieces only: pieces at the diagonal or the edge go element by
     * element); all of a thread's loads are issued before its LDS stores */
    const bool src_i = s.src.rs == 1;
    if (s.vsrc) {
        typedef TilePiece<VN> TP;
        T x[TP::PASSES][VN];
#pragma unroll
        for (int e = 0; e < TP::PASSES; ++e) {
            int li, lj;
            TP::at(tid, e, src_i, li, lj);
            const int64_t gi = i0 + li, gj = j0 + lj;
            const int64_t gie = src_i ? gi + VN - 1 : gi, gje = src_i ? gj : gj + VN - 1;
            const T *ptr = src + s.src.at(gi, gj);
            if (inside(gie, gje) && inside(gi, gj) && (!s.tri || gi >= gje)) {
                const V v = *(const V *)ptr;
#pragma unroll
                for (int c = 0; c < VN; ++c) x[e][c] = v[c];
            } else {
#pragma unroll
                for (int c = 0; c < VN; ++c)
                    x[e][c] = inside(src_i ? gi + c : gi, src_i ? gj : gj + c) ? ptr[c] : T(0);
            }
        }
#pragma unroll
        for (int e = 0; e < TP::PASSES; ++e) {
            int li, lj;
            TP::at(tid, e, src_i, li, lj);
#pragma unroll
            for (int c = 0; c < VN; ++c) tile[src_i ? li + c : li][src_i ? lj : lj + c] = x[e][c];
        }
    } else {
        typedef TilePiece<1> TP;
        T x[TP::PASSES];
#pragma unroll
        for (int e = 0; e < TP::PASSES; ++e) {
            int li, lj;
            TP::at(tid, e, src_i, li, lj);
            const int64_t gi = i0 + li, gj = j0 + lj;
            x[e] = inside(gi, gj) ? src[s.src.at(gi, gj)] : T(0);
        }
#pragma unroll
        for (int e = 0; e < TP::PASSES; ++e) {
            int li, lj;
            TP::at(tid, e, src_i, li, lj);
            tile[li][lj] = x[e];
        }
    }
    __syncthreads();

    /* write along the destination's contiguous direction, the same way */
    const bool dst_i = s.dst.rs == 1;
    if (s.vdst) {
        typedef TilePiece<VN> TP;
#pragma unroll
        for (int e = 0; e < TP::PASSES; ++e) {
            int li, lj;
            TP::at(tid, e, dst_i, li, lj);
            const int64_t gi = i0 + li, gj = j0 + lj;
            const int64_t gie = dst_i ? gi + VN - 1 : gi, gje = dst_i ? gj : gj + VN - 1;
            T *ptr = dst + s.dst.at(gi, gj);
            if (inside(gie, gje) && inside(gi, gj) && (!s.tri || gi >= gje)) {
                V v;
#pragma unroll
                for (int c = 0; c < VN; ++c) v[c] = tile[dst_i ? li + c : li][dst_i ? lj : lj + c];
                *(V *)ptr = v;
            } else {
#pragma unroll
                for (int c = 0; c < VN; ++c) {
                    const int64_t ei = dst_i ? gi + c : gi, ej = dst_i ? gj : gj + c;
                    if (inside(ei, ej)) ptr[c] = tile[dst_i ? li + c : li][dst_i ? lj : lj + c];
                }
            }
        }
    } else {
        typedef TilePiece<1> TP;
#pragma unroll
        for (int e = 0; e < TP::PASSES; ++e) {
            int li, lj;
            TP::at(tid, e, dst_i, li, lj);
            const int64_t gi = i0 + li, gj = j0 + lj;
            if (inside(gi, gj)) dst[s.dst.at(gi, gj)] = tile[li][lj];
        }
    }
}

/* Logical lower element (i, j) of the dense row-major array: A[i*lda + j]
 * for uplo = L, A[j*lda + i] (the upper triangle) for uplo = U. */
static RfpView dense_view(bool lower, int64_t lda)
{
    return lower ? RfpView{0, lda, 1} : RfpView{0, 1, lda};
}

static CopyParams make_copy(const RfpGeom &g, bool lower, int64_t lda, bool to_rfp)
{
    CopyParams p;
    const RfpView d = dense_view(lower, lda);
    const RfpView dv[3] = {d, d.sub(g.n1, 0), d.sub(g.n1, g.n1)};
    const RfpView rv[3] = {g.v11, g.v21, g.v22};
    const int64_t rows[3] = {g.n1, g.n2, g.n2};
    const int64_t cols[3] = {g.n1, g.n1, g.n2};
    const int tri[3] = {1, 0, 1};
    int64_t first = 0;
    p.nseg = 0;
    for (int k = 0; k < 3; ++k) {
        if (rows[k] <= 0 || cols[k] <= 0) continue;
        CopySeg &s = p.seg[p.nseg++];
        s.src = to_rfp ? dv[k] : rv[k];
        s.dst = to_rfp ? rv[k] : dv[k];
        s.rows = rows[k];
        s.cols = cols[k];
        s.tri = tri[k];
        s.tiles_r = (rows[k] + PT - 1) / PT;
        s.tiles_c = (cols[k] + PT - 1) / PT;
        s.first = first;
        first += s.tri ? s.tiles_r * (s.tiles_r + 1) / 2 : s.tiles_r * s.tiles_c;
    }
    for (int k = p.nseg; k < 3; ++k) p.seg[k].first = first;
    return p;
}

static int64_t copy_blocks(const CopyParams &p)
{
    if (p.nseg == 0) return 0;
    const CopySeg &l = p.seg[p.nseg - 1];
    return l.first + (l.tri ? l.tiles_r * (l.tiles_r + 1) / 2 : l.tiles_r * l.tiles_c);
}

/* a view's 16-byte pieces along its contiguous direction are aligned */
template <typename T>
static int vec_ok(const T *base, const RfpView &v)
{
    constexpr int64_t VN = 16 / sizeof(T);
    const int64_t ld = v.rs == 1 ? v.cs : v.rs;
    return ((uintptr_t)base % 16 == 0 && v.off % VN == 0 && ld % VN == 0) ? 1 : 0;
}

template <typename T>
static hipError_t rfp_copy(const T *src, T *dst, CopyParams p, hipStream_t s)
{
    const int64_t nb = copy_blocks(p);
    if (nb <= 0) return hipSuccess;
    for (int k = 0; k < p.nseg; ++k) {
        p.seg[k].vsrc = vec_ok(src, p.seg[k].src);
        p.seg[k].vdst = vec_ok(dst, p.seg[k].dst);
    }
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((rfp_copy_kernel<T>), dim3((unsigned)nb), dim3(256), 0, s, src, dst, p);
    return hipGetLastError();
}

template <typename T>
__global__ void add_diag_kernel(T *arf, int64_t n, int64_t n1, RfpView v11, RfpView v22, T value)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t idx = i < n1 ? v11.at(i, i) : v22.at(i - n1, i - n1);
    arf[idx] += value;
}

template <typename T>
static hipError_t add_diag(const RfpGeom &g, T value, T *arf, hipStream_t s)
{
    if (g.n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((g.n + 255) / 256);
    hipLaunchKernelGGL((add_diag_kernel<T>), dim3(blocks), dim3(256), 0, s, arf, g.n, g.n1, g.v11, g.v22, value);
    return hipGetLastError();
}

__global__ void fill_zero_kernel(uint32_t *p, int64_t words)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0u;
}

__global__ void set_int_kernel(int *p, int v) { *p = v; }

}  // namespace


hipError_t launch_trttf_f32(const RfpGeom &g, bool lower, const float *A, int64_t lda, float *arf, hipStream_t s)
{ return rfp_copy<float>(A, arf, make_copy(g, lower, lda, true), s); }
hipError_t launch_trttf_f64(const RfpGeom &g, bool lower, const double *A, int64_t lda, double *arf, hipStream_t s)
{ return rfp_copy<double>(A, arf, make_copy(g, lower, lda, true), s); }
hipError_t launch_tfttr_f32(const RfpGeom &g, bool lower, const float *arf, float *A, int64_t lda, hipStream_t s)
{ return rfp_copy<float>(arf, A, make_copy(g, lower, lda, false), s); }
hipError_t launch_tfttr_f64(const RfpGeom &g, bool lower, const double *arf, double *A, int64_t lda, hipStream_t s)
{ return rfp_copy<double>(arf, A, make_copy(g, lower, lda, false), s); }

hipError_t launch_add_diag_f32(const RfpGeom &g, float value, float *arf, hipStream_t s)
{ return add_diag<float>(g, value, arf, s); }
hipError_t launch_add_diag_f64(const RfpGeom &g, double value, double *arf, hipStream_t s)
{ return add_diag<double>(g, value, arf, s); }

/* Byte counts here are always multiples of 4 (float / double arrays). */
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s)
{
    const int64_t words = (int64_t)(bytes / 4);
    if (words <= 0) return hipSuccess;
    int64_t blocks = (words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t *)p, words);
    return hipGetLastError();
}

hipError_t launch_set_int(int *p, int value, hipStream_t s)
{
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, value);
    return hipGetLastError();
}
