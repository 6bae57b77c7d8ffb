/*
 * tile_stage.h — staging of 128 x 128 FP32 operand tiles from strided RFP
 * views through registers into LDS, shared by the triangular-solve kernels
 * (trsv_block.hip) and the panel TRSM of the factorisation (trsm_panel.hip).
 */
#pragma once

#include "rfp_internal.h"

typedef float f32x4v __attribute__((ext_vector_type(4)));
/* 16-byte vector that only assumes element alignment (RFP leading dimensions
 * such as n + 1 leave rows unaligned) */
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

#define TILE_LF 128

/* Workgroup barrier that orders LDS accesses only.  __syncthreads() also
 * orders global memory, which makes every wave wait for all of its
 * outstanding global loads first — including operands prefetched for later
 * steps.  With LDS-only fences the waves wait just for their LDS traffic and
 * the prefetches stay in flight across the barrier.  Not for ordering
 * global-memory communication between the waves of a workgroup. */
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

/* The 128 x 128 operand A(i, k) = base[off + i*rs + k*cs] of one step (rows /
 * columns at or beyond vi / vk read as zero). */
struct Blk {
    const float *base;
    int64_t off, rs, cs;
    int vi, vk;
};

/* A thread's 64 elements of a staged operand as 16 vectors of 4 consecutive
 * elements along the contiguous direction of the source (eight rows /
 * columns apart). */
struct Staged {
    f32x4v v[16];

    /* WHOLE: a straight-line path for whole tiles (fewer registers in
     * kernels that stage one tile at a time; kernels that unroll many
     * staging sites keep the single masked path) */
    template <bool WHOLE = true>
    __device__ __forceinline__ void load(const Blk &b, int tid)
    {
        const int c4 = (tid % 32) * 4, o0 = tid / 32;
        const bool kc = b.cs == 1;
        /* kc: lanes run along k (c4 = k), o = i; else along i (c4 = i), o = k
         * (values, not member lvalues, in the selects: no stack copy of b) */
        const int vi = b.vi, vk = b.vk;
        const int64_t rs = b.rs, cs = b.cs;
        const int vc = kc ? vk : vi, vo = kc ? vi : vk;
        const int64_t so = kc ? rs : cs;
        const float *p = b.base + b.off + c4 + (int64_t)o0 * so;
        const int64_t step = 8 * so;
        if (WHOLE && vi == TILE_LF && vk == TILE_LF) {
            /* whole tile (the common case): straight-line loads */
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const f32x4u x = *(const f32x4u *)(p + e * step);
                v[e] = f32x4v{x[0], x[1], x[2], x[3]};
            }
            return;
        }
        const bool full = c4 + 4 <= vc;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int o = o0 + 8 * e;
            const float *q = p + e * step;
            if (o < vo && full) {
                const f32x4u x = *(const f32x4u *)q;
                v[e] = f32x4v{x[0], x[1], x[2], x[3]};
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) v[e][c] = (o < vo && c4 + c < vc) ? q[c] : 0.f;
            }
        }
    }
    /* into LDS along the source's contiguous direction: As[i*pitch + k] when
     * k is contiguous (cs == 1), else As[k*pitch + i]; 16-byte stores, so the
     * pitch is a multiple of 4 */
    __device__ __forceinline__ void store(float *As, int pitch, int tid) const
    {
        const int c4 = (tid % 32) * 4, o0 = tid / 32;
#pragma unroll
        for (int e = 0; e < 16; ++e) *(f32x4v *)(As + (o0 + 8 * e) * pitch + c4) = v[e];
    }
};
