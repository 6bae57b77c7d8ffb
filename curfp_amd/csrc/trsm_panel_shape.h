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
constexpr int BF16_MAX_ROWS = 16384;                /* BF16 modes: taller panels keep the recursion */
}  // namespace panel_shape
