/*
 * panel_replay.cpp — host replay of every global-memory access of the panel
 * TRSM kernel (curfp_amd/csrc/trsm_panel.hip, trsm_panel_kernel, lines
 * 145-352, with the operand staging of tile_stage.h, Staged::load) for given
 * launch geometries, each access checked against the buffer it belongs to.
 * Built with g++ and run by tests/test_panel_addresses.py; no GPU involved.
 *
 * The launch shape (panel_shape::*), the RFP geometry (rfp_geom), the BF16
 * shadow geometry (shadow_geom) and the recursion's split rule (leaf_split)
 * come from the library's own headers; the per-thread index arithmetic is a
 * replica of the kernel's, with the kernel's line numbers.
 *
 *   panel_replay probe [rows]
 *       the stand-alone geometry of an earlier instrumented probe: L a
 *       512 x 512 row view at 0 and X = rows 512.. of the same row-major
 *       array with ld = 8192, mb = 32768 (1024 workgroups), the array holding
 *       `rows` rows of 8192 (default 512 + 32768, the exact extent); every
 *       workgroup replayed
 *   panel_replay chain n transr uplo bf16 [all]
 *       every panel launch of spftrf at order n (Factor::potrf / potrf2 /
 *       trsm_rt of rfp_drivers.hip), for the FP32 (bf16 = 0), BF16 (1) or
 *       split-BF16 (3) modes (BF16: shadow writes; FP32 and BF16: no panel
 *       launch inside a block of up to two leaves, which is one pair
 *       launch);
 *       workgroups 0, 1 and the last two of each launch (every address of a
 *       workgroup is its first row r0 = 32 * blockIdx.x times the row stride
 *       plus a pattern that depends on blockIdx.x only through the row count
 *       of the last, ragged workgroup), or all workgroups with `all`
 *   panel_replay chain64 n transr uplo [all]
 *       the same for the FP64 panel kernel (trsm_panel64.hip) over every
 *       panel launch of dpftrf at order n (64-row leaves, panels of at most
 *       four leaves)
 *
 * Prints one summary line and exits 0, or prints the first violation and
 * exits 1.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rfp_internal.h"
#include "trsm_panel_shape.h"

using namespace panel_shape;

namespace {

struct Bounds {
    const char *name;
    int64_t lo, hi;   /* valid element offsets [lo, hi) */
};

struct Launch {
    RfpView X, L, S;     /* S: shadow view of X (bf16 only) */
    int64_t mb, m;
    int64_t inv_slot;    /* first leaf-inverse slot */
};

struct Replay {
    Bounds arr;          /* the array X and L live in */
    Bounds inv;          /* leaf-inverse slots (element offsets from the workspace start) */
    Bounds shadow;       /* BF16 shadow (elements) */
    bool bf16 = false;
    bool all_wg = false;
    int64_t checked = 0, launches = 0, lo_seen = INT64_MAX, hi_seen = -1;
    std::string where;

    [[noreturn]] void fail(const std::string &what, int64_t v, const Bounds &b)
    {
        std::printf("VIOLATION %s: %s offset %lld outside [%lld, %lld) (%s)\n", where.c_str(), b.name,
                    (long long)v, (long long)b.lo, (long long)b.hi, what.c_str());
        std::exit(1);
    }
    void at(const Bounds &b, int64_t v, const char *what)
    {
        if (v < b.lo || v >= b.hi) fail(what, v, b);
        ++checked;
        if (&b == &arr) {
            lo_seen = std::min(lo_seen, v);
            hi_seen = std::max(hi_seen, v);
        }
    }
    void coord(bool ok, const char *what, int64_t i, int64_t j)
    {
        if (!ok) {
            std::printf("VIOLATION %s: %s at view element (%lld, %lld)\n", where.c_str(), what, (long long)i,
                        (long long)j);
            std::exit(1);
        }
    }

    /* tile_stage.h Staged::load: piece e of thread tid covers 4 elements
     * along the contiguous direction, c4 = (tid % 32) * 4, o = tid / 32 + 8 e;
     * element c is read when o < vo and c4 + c < vc (the whole-tile path reads
     * the same elements when vi = vk = 128). */
    template <typename F>
    void stage(int64_t off, int64_t rs, int64_t cs, int vi, int vk, F &&elem)
    {
        const bool kc = cs == 1;
        const int vc = kc ? vk : vi, vo = kc ? vi : vk;
        for (int tid = 0; tid < THREADS; ++tid)
            for (int e = 0; e < 16; ++e) {
                const int c4 = (tid % 32) * 4, o = tid / 32 + 8 * e;
                for (int c = 0; c < 4; ++c) {
                    if (!(o < vo && c4 + c < vc)) continue;
                    const int i = kc ? o : c4 + c, k = kc ? c4 + c : o;
                    elem(i, k, off + (int64_t)i * rs + (int64_t)k * cs);
                }
            }
    }

    void launch(const Launch &p)
    {
        if (p.mb <= 0 || p.m <= 0) return;
        ++launches;
        const int m = (int)p.m, mb = (int)p.mb;
        const int nq = (m + LF - 1) / LF, wm = nq * LF, nsteps = nq * (nq + 1) / 2;
        const int64_t grid = (p.mb + RM - 1) / RM;
        /* operand tiles (trsm_panel.hip:156-167), the same for every workgroup:
         * step j -> leaf q, tile column pp (pp == q: the inverse of leaf q) */
        for (int j = 0; j < nsteps; ++j) {
            int q = 0, pp = j;
            while (pp > q) { pp -= q + 1; ++q; }
            const int mq = std::min(LF, m - q * LF);
            if (pp == q) {
                const int64_t base = (p.inv_slot + q) * (int64_t)LF * LF;
                stage(0, LF, 1, mq, mq, [&](int i, int k, int64_t a) {
                    coord(i < mq && k < mq, "inverse tile", i, k);
                    at(inv, base + a, "leaf inverse");
                });
            } else {
                const RfpView v = p.L.sub((int64_t)q * LF, (int64_t)pp * LF);
                stage(v.off, v.rs, v.cs, mq, LF, [&](int i, int k, int64_t a) {
                    const int64_t li = (int64_t)q * LF + i, lk = (int64_t)pp * LF + k;
                    coord(li < m && lk < li, "L tile outside the strict lower triangle", li, lk);
                    at(arr, a, "L tile");
                });
            }
        }
        std::vector<int64_t> wgs;
        if (all_wg) {
            for (int64_t b = 0; b < grid; ++b) wgs.push_back(b);
        } else {
            for (int64_t b : {(int64_t)0, (int64_t)1, grid - 2, grid - 1})
                if (b >= 0 && b < grid && (wgs.empty() || wgs.back() < b)) wgs.push_back(b);
        }
        const bool xc = p.X.cs == 1;
        for (int64_t b : wgs) {
            const int64_t r0 = b * RM;
            const int rows = (int)std::min<int64_t>(RM, p.mb - r0);
            /* slab load (:184-217) and store (:299-339): the same pieces */
            for (int tid = 0; tid < THREADS; ++tid)
                for (int e = 0; e < SL; ++e) {
                    const int idx = tid + THREADS * e;
                    int r, c;
                    if (xc) { r = idx / (MAXQ * LF / 4); c = (idx % (MAXQ * LF / 4)) * 4; }
                    else    { r = (idx % (RM / 4)) * 4;  c = idx / (RM / 4); }
                    coord(r >= 0 && r < RM && c >= 0 && c + (xc ? 3 : 0) < XP, "slab piece outside LDS", r, c);
                    if (xc) {
                        if (!(r < rows && c < m)) continue;
                        const int cnt = c + 4 <= m ? 4 : m - c;
                        for (int q = 0; q < cnt; ++q) {
                            coord(r0 + r < mb && c + q < m, "X element", r0 + r, c + q);
                            at(arr, p.X.at(r0 + r, c + q), "X slab");
                        }
                    } else {
                        if (!(c < m && r < rows)) continue;
                        const int cnt = r + 4 <= rows ? 4 : rows - r;
                        for (int q = 0; q < cnt; ++q) {
                            coord(r0 + r + q < mb && c < m, "X element", r0 + r + q, c);
                            at(arr, p.X.at(r0 + r + q, c), "X slab");
                        }
                    }
                }
            /* shadow writes (the end of trsm_panel_kernel): eight columns per
             * thread and store along 16-byte aligned k-contiguous shadow rows,
             * else one element per thread and store */
            if (bf16 && p.S.cs == 1 && p.S.rs % 8 == 0 && p.S.off % 8 == 0) {
                const int CH = MAXQ * LF / 8;
                for (int tid = 0; tid < THREADS; ++tid)
                    for (int e = 0; e < RM * CH / THREADS; ++e) {
                        const int idx = tid + THREADS * e, r = idx / CH, c = (idx % CH) * 8;
                        if (r >= rows || c >= m) continue;
                        coord(r < RM && c + 7 < XP, "shadow source outside LDS", r, c);
                        const int cnt = c + 8 <= m ? 8 : m - c;
                        for (int q = 0; q < cnt; ++q) at(shadow, p.S.at(r0 + r, c + q), "BF16 shadow");
                    }
            } else if (bf16) {
                for (int tid = 0; tid < THREADS; ++tid)
                    for (int idx = tid; idx < RM * wm; idx += THREADS) {
                        const int r = idx / wm, c = idx % wm;
                        if (r >= rows || c >= m) continue;
                        coord(r < RM && c < XP, "shadow source outside LDS", r, c);
                        at(shadow, p.S.at(r0 + r, c), "BF16 shadow");
                    }
            }
        }
    }
};

/* FP64 panel (curfp_amd/csrc/trsm_panel64.hip, trsm_panel64_kernel): the
 * tile and slab element maps are the kernel's own (panel64_shape::tile_elem,
 * slab_elem); the masks are replicas of tile_load and the slab loops. */
struct Replay64 : Replay {
    void launch(const Launch &p)
    {
        namespace P = panel64_shape;
        if (p.mb <= 0 || p.m <= 0) return;
        ++launches;
        const int m = (int)p.m, mb = (int)p.mb;
        const int nq = (m + P::LF - 1) / P::LF, nsteps = nq * (nq + 1) / 2;
        const int64_t grid = (p.mb + P::RM - 1) / P::RM;
        for (int j = 0; j < nsteps; ++j) {
            int q = 0, pp = j;                           /* step_of */
            while (pp > q) { pp -= q + 1; ++q; }
            const int mq = std::min(P::LF, m - q * P::LF);
            const RfpView t = p.L.sub((int64_t)q * P::LF, (int64_t)pp * P::LF);
            for (int tid = 0; tid < P::THREADS; ++tid)
                for (int e = 0; e < P::TE; ++e) {
                    int c, o;
                    P::tile_elem(tid, e, c, o);
                    coord(c >= 0 && c < P::LF && o >= 0 && o < P::LF, "tile element outside the tile", o, c);
                    if (pp == q) {
                        if (!(o < mq && c < mq)) continue;
                        at(inv, (p.inv_slot + q) * (int64_t)P::LF * P::LF + o * P::LF + c, "leaf inverse");
                    } else {
                        const bool kc = p.L.cs == 1;
                        const int n = kc ? o : c, k = kc ? c : o;
                        if (!(n < mq)) continue;
                        const int64_t li = (int64_t)q * P::LF + n, lk = (int64_t)pp * P::LF + k;
                        coord(li < m && lk < li, "L tile outside the strict lower triangle", li, lk);
                        at(arr, t.at(n, k), "L tile");
                    }
                }
        }
        std::vector<int64_t> wgs;
        if (all_wg) {
            for (int64_t b = 0; b < grid; ++b) wgs.push_back(b);
        } else {
            for (int64_t b : {(int64_t)0, (int64_t)1, grid - 2, grid - 1})
                if (b >= 0 && b < grid && (wgs.empty() || wgs.back() < b)) wgs.push_back(b);
        }
        const bool xc = p.X.cs == 1;
        for (int64_t b : wgs) {
            const int64_t r0 = b * P::RM;
            /* slab load and store: the same elements */
            for (int tid = 0; tid < P::THREADS; ++tid)
                for (int e = 0; e < P::SE; ++e) {
                    int r, c;
                    P::slab_elem(tid, e, xc, r, c);
                    coord(r >= 0 && r < P::RM && c >= 0 && c < P::MAXQ * P::LF, "slab element outside LDS", r, c);
                    if (!(r0 + r < mb && c < m)) continue;
                    at(arr, p.X.at(r0 + r, c), "X slab");
                }
        }
    }
};

/* the panel launches of rfp_drivers.hip Factor<double> (64-row leaves, every
 * trsm_rt with m <= 256 is one panel launch) */
struct Chain64 {
    Replay64 &rp;
    void trsm_rt(RfpView B, int64_t mb, RfpView L, int64_t m, int64_t slot)
    {
        const int64_t leaf = panel64_shape::LF;
        if (m <= 0 || mb <= 0) return;
        if (m <= panel64_shape::MAXQ * leaf) {
            rp.launch(Launch{B, L, RfpView{0, 0, 0}, mb, m, slot});
            return;
        }
        const int64_t m1 = leaf_split(m, leaf);
        trsm_rt(B, mb, L, m1, slot);
        trsm_rt(B.sub(0, m1), mb, L.sub(m1, m1), m - m1, slot + m1 / leaf);
    }
    void potrf2(RfpView V11, RfpView V21, RfpView V22, int64_t m1, int64_t m2, int64_t slot1, int64_t slot2)
    {
        potrf(V11, m1, slot1);
        trsm_rt(V21, m2, V11, m1, slot1);
        potrf(V22, m2, slot2);
    }
    void potrf(RfpView V, int64_t m, int64_t slot)
    {
        const int64_t leaf = panel64_shape::LF;
        if (m <= leaf) return;
        const int64_t m1 = leaf_split(m, leaf);
        potrf2(V, V.sub(m1, 0), V.sub(m1, m1), m1, m - m1, slot, slot + m1 / leaf);
    }
};

/* the panel launches of rfp_drivers.hip Factor<float>::potrf2 / potrf /
 * trsm_rt (the GEMM and leaf steps in between launch other kernels) */
struct Chain {
    Replay &rp;
    int64_t leaf;
    bool bf16;
    bool pair;   /* FP32 / BF16: blocks of up to two leaves are one pair launch */

    void trsm_rt(RfpView B, RfpView Bs, int64_t mb, RfpView L, RfpView Ls, int64_t m, int64_t slot)
    {
        if (m <= 0 || mb <= 0) return;
        if (m <= MAXQ * LF) {
            rp.launch(Launch{B, L, Bs, mb, m, slot});
            return;
        }
        if (m <= leaf) return;
        const int64_t m1 = leaf_split(m, leaf);
        trsm_rt(B, Bs, mb, L, Ls, m1, slot);
        trsm_rt(B.sub(0, m1), Bs.sub(0, m1), mb, L.sub(m1, m1), Ls.sub(m1, m1), m - m1, slot + m1 / leaf);
    }
    void potrf2(RfpView V11, RfpView S11, RfpView V21, RfpView S21, RfpView V22, RfpView S22, int64_t m1, int64_t m2,
                int64_t slot1, int64_t slot2)
    {
        potrf(V11, S11, m1, slot1);
        trsm_rt(V21, S21, m2, V11, S11, m1, slot1);
        potrf(V22, S22, m2, slot2);
    }
    void potrf(RfpView V, RfpView S, int64_t m, int64_t slot)
    {
        if (m <= leaf || (pair && m <= 2 * leaf)) return;
        const int64_t m1 = leaf_split(m, leaf);
        potrf2(V, S, V.sub(m1, 0), S.sub(m1, 0), V.sub(m1, m1), S.sub(m1, m1), m1, m - m1, slot, slot + m1 / leaf);
    }
};

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: panel_replay probe [rows] | chain n transr uplo bf16 [all] | chain64 n transr uplo [all]\n");
        return 2;
    }
    Replay rp;
    if (!std::strcmp(argv[1], "probe")) {
        const int64_t ld = 8192, m = 512, mb = 32768;
        const int64_t rows = argc > 2 ? std::atoll(argv[2]) : m + mb;
        rp.arr = Bounds{"array", 0, rows * ld};
        rp.inv = Bounds{"inverses", 0, MAXQ * (int64_t)LF * LF};
        rp.shadow = Bounds{"shadow", 0, 0};
        rp.all_wg = true;
        rp.where = "probe mb=32768 m=512 ld=8192";
        rp.launch(Launch{RfpView{m * ld, ld, 1}, RfpView{0, ld, 1}, RfpView{0, 0, 0}, mb, m, 0});
    } else if (!std::strcmp(argv[1], "chain") && argc >= 6) {
        const int64_t n = std::atoll(argv[2]);
        const bool tn = argv[3][0] == 'N', lower = argv[4][0] == 'L';
        const int mode = std::atoi(argv[5]);   /* 0 FP32, 1 BF16, 3 split BF16 */
        rp.bf16 = mode != 0;
        rp.all_wg = argc > 6 && !std::strcmp(argv[6], "all");
        const RfpGeom g = rfp_geom(n, tn, lower);
        const ShadowGeom sg = shadow_geom(g);
        const int64_t leaf = LF;
        const int64_t s1 = (g.n1 + leaf - 1) / leaf, s2 = (g.n2 + leaf - 1) / leaf;
        rp.arr = Bounds{"RFP array", 0, n * (n + 1) / 2};
        rp.inv = Bounds{"inverses", 0, (s1 + s2) * leaf * leaf};
        rp.shadow = Bounds{"shadow", 0, rp.bf16 ? sg.elems : 0};
        rp.where = "chain n=" + std::to_string(n) + " " + argv[3] + argv[4] + (rp.bf16 ? " bf16" : " fp32");
        Chain ch{rp, leaf, rp.bf16, mode != 3};
        ch.potrf2(g.v11, sg.s11, g.v21, sg.s21, g.v22, sg.s22, g.n1, g.n2, 0, s1);
    } else if (!std::strcmp(argv[1], "chain64") && argc >= 5) {
        const int64_t n = std::atoll(argv[2]);
        const bool tn = argv[3][0] == 'N', lower = argv[4][0] == 'L';
        Replay64 r64;
        r64.all_wg = argc > 5 && !std::strcmp(argv[5], "all");
        const RfpGeom g = rfp_geom(n, tn, lower);
        const int64_t leaf = panel64_shape::LF;
        const int64_t s1 = (g.n1 + leaf - 1) / leaf, s2 = (g.n2 + leaf - 1) / leaf;
        r64.arr = Bounds{"RFP array", 0, n * (n + 1) / 2};
        r64.inv = Bounds{"inverses", 0, (s1 + s2) * leaf * leaf};
        r64.shadow = Bounds{"shadow", 0, 0};
        r64.where = "chain64 n=" + std::to_string(n) + " " + argv[3] + argv[4];
        Chain64 ch{r64};
        ch.potrf2(g.v11, g.v21, g.v22, g.n1, g.n2, 0, s1);
        rp = r64;
    } else {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    std::printf("ok %s: %lld launches, %lld accesses, array offsets [%lld, %lld] within [0, %lld)\n",
                rp.where.c_str(), (long long)rp.launches, (long long)rp.checked, (long long)rp.lo_seen,
                (long long)rp.hi_seen, (long long)rp.arr.hi);
    return 0;
}
