/*
 * mfma_epilogue.h — writing one 32 x 32 block of v_mfma_f32_32x32x* output
 * to a strided FP32 view, shared by the GEMM kernels.
 *
 * Accumulator element 4 g + e of lane l holds row 8 g + 4 (l / 32) + e and
 * column l % 32 of the block, so a lane owns, per g, a run of 4 consecutive
 * rows of one column.  When the view's rows are contiguous (a column view of
 * the RFP array, or a column-major partial) each run is written with one
 * 16-byte store (and read with one 16-byte load when beta != 0) instead of
 * four scattered 4-byte accesses; runs that cross the edge of C or the
 * diagonal of a lower-triangular output fall back to element stores.
 */
#pragma once

#include <stdint.h>

typedef float f32x4run __attribute__((ext_vector_type(4), aligned(4)));

/* C(row0 + i, col0 + j) = alpha acc + beta C for the block's elements with
 * i < M - row0, j < N - col0 (and row >= column when lower) */
template <typename ACC>
__device__ __forceinline__ void store_block32(const ACC &acc, float *c, int64_t c_off, int64_t c_rs, int64_t c_cs,
                                              int64_t row0, int64_t col0, int64_t M, int64_t N, bool lower,
                                              float alpha, float beta, int lane)
{
    const int64_t gj = col0 + (lane & 31);
    const bool use_beta = beta != 0.f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int64_t gi0 = row0 + 8 * g + 4 * (lane >> 5);
        if (c_rs == 1 && gi0 + 3 < M && gj < N && (!lower || gi0 >= gj)) {
            float *cp = c + c_off + gi0 + gj * c_cs;
            f32x4run v = {alpha * acc[4 * g], alpha * acc[4 * g + 1], alpha * acc[4 * g + 2], alpha * acc[4 * g + 3]};
            if (use_beta) v = v + beta * *(const f32x4run *)cp;
            *(f32x4run *)cp = v;
            continue;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t gi = gi0 + e;
            if (gi < M && gj < N && (!lower || gi >= gj)) {
                float *cp = c + c_off + gi * c_rs + gj * c_cs;
                float val = alpha * acc[4 * g + e];
                if (use_beta) val += beta * (*cp);
                *cp = val;
            }
        }
    }
}

/* The same for an accumulator that holds the block transposed (the MFMA
 * operands exchanged): element 4 g + e of lane l is row l % 32 and column
 * 8 g + 4 (l / 32) + e, so a lane owns runs of 4 consecutive columns of one
 * row — one 16-byte access when the view's columns are contiguous (a row
 * view). */
template <typename ACC>
__device__ __forceinline__ void store_block32_t(const ACC &acc, float *c, int64_t c_off, int64_t c_rs, int64_t c_cs,
                                                int64_t row0, int64_t col0, int64_t M, int64_t N, bool lower,
                                                float alpha, float beta, int lane)
{
    const int64_t gi = row0 + (lane & 31);
    const bool use_beta = beta != 0.f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int64_t gj0 = col0 + 8 * g + 4 * (lane >> 5);
        if (c_cs == 1 && gi < M && gj0 + 3 < N && (!lower || gi >= gj0 + 3)) {
            float *cp = c + c_off + gi * c_rs + gj0;
            f32x4run v = {alpha * acc[4 * g], alpha * acc[4 * g + 1], alpha * acc[4 * g + 2], alpha * acc[4 * g + 3]};
            if (use_beta) v = v + beta * *(const f32x4run *)cp;
            *(f32x4run *)cp = v;
            continue;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t gj = gj0 + e;
            if (gi < M && gj < N && (!lower || gi >= gj)) {
                float *cp = c + c_off + gi * c_rs + gj * c_cs;
                float val = alpha * acc[4 * g + e];
                if (use_beta) val += beta * (*cp);
                *cp = val;
            }
        }
    }
}
