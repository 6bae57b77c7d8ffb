/*
 * trsm_panel_shape.h — launch shape of the panel TRSM (trsm_panel.hip),
 * shared by the kernel, the factorisation driver that launches it and the
 * host replay of its global-memory addresses (tests/panel_replay.cpp).
 */
#pragma once

namespace panel_shape {
constexpr int LF = 128;                             /* rows of a diagonal leaf */
constexpr int MAXQ = 4;                             /* leaves per panel */
constexpr int RM = 32;                              /* rows of X per workgroup */
constexpr int THREADS = 256;                        /* threads per workgroup */
constexpr int XP = MAXQ * LF + 4;                   /* LDS pitch of the slab (== 4 mod 64) */
constexpr int SP = LF + 4;                          /* LDS pitch of a staged tile */
constexpr int SL = RM * MAXQ * LF / 4 / THREADS;    /* 16-byte slab pieces per thread */
}  // namespace panel_shape

/* FP64 panel (trsm_panel64.hip): 64-row leaves */
namespace panel64_shape {
constexpr int LF = 64;                              /* rows of a diagonal leaf */
constexpr int MAXQ = 4;                             /* leaves per panel */
constexpr int RM = 32;                              /* rows of X per workgroup */
constexpr int THREADS = 256;                        /* threads per workgroup */
constexpr int XP = MAXQ * LF + 1;                   /* LDS pitch of the slab (doubles) */
constexpr int TP = LF + 1;                          /* LDS pitch of a staged tile (doubles) */
constexpr int TE = LF * LF / THREADS;               /* tile elements per thread */
constexpr int SE = RM * MAXQ * LF / THREADS;        /* slab elements per thread */

/* Tile element e of thread tid: c runs along the source's contiguous
 * direction (64 consecutive lanes), o across it. */
inline __host__ __device__ void tile_elem(int tid, int e, int &c, int &o)
{
    c = tid % LF;
    o = tid / LF + (THREADS / LF) * e;
}

/* Slab element e of thread tid as (row r, column c) of the workgroup's
 * RM x MAXQ*LF slab; `along_c`: the view is contiguous along columns. */
inline __host__ __device__ void slab_elem(int tid, int e, bool along_c, int &r, int &c)
{
    if (along_c) {
        c = tid % 64 + 64 * (e % 4);
        r = tid / 64 + 4 * (e / 4);
    } else {
        r = tid % RM;
        c = tid / RM + (THREADS / RM) * e;
    }
}
}  // namespace panel64_shape
