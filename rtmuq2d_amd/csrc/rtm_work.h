/*
 * rtm_work.h -- how one simulation's stepped region 4 <= i < X-4,
 * 4 <= k < Z-4 is split into work items for the step kernels.  Shared by
 * the library and the CPU test harness so both walk the same items.
 *
 * Tiles are 64 x TZ (TZ = 4*rz, or 8*rz for IN_TILES8) from (4, 4).  Each
 * tile has a class: which CPML strips its points touch.  The interior
 * (class-0 tiles, whose points need no CPML term) is handled by one of two
 * kernels:
 *   IN_TILES    every tile goes to the tile kernel;
 *   IN_TILES8   the same with workgroups of 8 wavefronts, each owning rz
 *               rows (tiles of 64 x 8rz);
 *   IN_COLUMNS  class-0 runs of two adjacent tile columns with the same
 *               rows become one 128-column item, a lone column a 64-column
 *               item, in segments of at most seg_rows rows (rtm_column.h);
 *   IN_MIXED    class-0 tiles stay 64 x TZ and every other tile is cut into
 *               64 x 4 tiles (class | T_SMALL, the class of the small tile's
 *               own rows), all in one tile launch: the strip tiles' longer
 *               chain runs on less work per workgroup;
 *   IN_PAIRS    the tiles of the two-step region D2 (rtm_pair.h) become
 *               pair items of up to 112 columns in segments of at most
 *               seg_rows rows, stepped two steps per launch; every other
 *               tile stays a tile.
 */
#ifndef RTM_WORK_H
#define RTM_WORK_H

#include <algorithm>
#include <vector>

#include "rtm_pair.h"

namespace rtm {

enum Interior : int { IN_TILES = 0, IN_COLUMNS = 1, IN_MIXED = 2, IN_TILES8 = 3, IN_PAIRS = 4 };

/* rows of a tile of interior kind `interior` at rz rows per thread */
inline int tile_rows(int rz, int interior) { return (interior == IN_TILES8 ? 8 : 4) * rz; }

struct WorkItem {
    int a, b, c, d;   /* tile {sim, x0, z0, class}; interior {sim(+half), x0, zA, zB} */
};

inline void build_work(std::vector<WorkItem> &tiles, std::vector<WorkItem> &inner, int sim, int X, int Z, int TL,
                       int TR, int TU, int TD, int rz, int interior, int seg_rows)
{
    const int TZ = tile_rows(rz, interior);
    SimDesc<float> d{};
    d.X = X; d.Z = Z; d.TL = TL; d.TR = TR; d.TU = TU; d.TD = TD;
    int seg = 0;
    if (interior == IN_COLUMNS) seg = seg_rows > 0 ? (seg_rows + 3) / 4 * 4 : 1 << 30;
    PairRegion D2 = interior == IN_PAIRS ? pair_region(X, Z, TL, TR, TU, TD, TZ) : PairRegion{0, 0, 0, 0};
    /* the pair items' bounds hold at the region's corners, hence for all */
    if (!D2.empty() && !(pair_item_ok(1, D2.x0, D2.z0, D2.z1, X, Z) &&
                         pair_item_ok((D2.x1 - D2.x0 - 1) % kPairCols / 2 + 1,
                                      D2.x0 + (D2.x1 - D2.x0 - 1) / kPairCols * kPairCols, D2.z0, D2.z1, X, Z)))
        D2 = PairRegion{0, 0, 0, 0};
    auto in_d2 = [&](int x0, int z0) { return x0 >= D2.x0 && x0 < D2.x1 && z0 >= D2.z0 && z0 < D2.z1; };
    /* Items are listed x fastest (a row of tiles, then the next row), so the
     * workgroups an XCD runs at the same time cover whole grid rows: every
     * DRAM page opened for a row is read through, not 256 bytes of it. */
    struct Run { int x0, zA, zB; };
    std::vector<Run> runs;
    for (int x0 = kHalf; x0 < X - kHalf; x0 += 64) {
        int zA = -1, zB = -1;
        for (int z0 = kHalf; z0 < Z - kHalf; z0 += TZ)
            if (seg && tile_class(d, x0, z0, TZ) == 0) {
                /* all 64 x TZ points lie in the stepped region (the strips
                 * are at least 9 points thick) */
                if (zA < 0) zA = z0;
                zB = z0 + TZ;
            }
        runs.push_back({x0, zA, zB});
    }
    for (int z0 = kHalf; z0 < Z - kHalf; z0 += TZ) {
        int t = 0;
        for (int x0 = kHalf; x0 < X - kHalf; x0 += 64, t++) {
            const Run &r = runs[t];
            if (r.zA >= 0 && z0 >= r.zA && z0 < r.zB) continue;
            if (!D2.empty() && in_d2(x0, z0)) continue;
            const int cls = tile_class(d, x0, z0, TZ);
            /* mixed lists cut every strip tile into 64 x 4 tiles; the other
             * four-row-group lists cut corner tiles (corner_tiles_small) */
            if ((interior == IN_MIXED && cls != 0) ||
                (cls == (T_XSTRIP | T_ZSTRIP) && interior != IN_TILES8 && corner_tiles_small(rz, 4))) {
                for (int zs = z0; zs < z0 + TZ && zs < Z - kHalf; zs += 4)
                    tiles.push_back({sim, x0, zs, tile_class(d, x0, zs, 4) | T_SMALL});
                continue;
            }
            tiles.push_back({sim, x0, z0, cls});
        }
    }
    /* interior items {x0, zA, zB, half}: one run per tile column, two
     * adjacent columns with the same rows paired into 128 columns */
    struct Col { int x0, zA, zB, half; };
    std::vector<Col> cols;
    for (size_t t = 0; t < runs.size(); t++) {
        const Run &r = runs[t];
        if (r.zA < 0) continue;
        const bool pair = t + 1 < runs.size() && runs[t + 1].zA == r.zA
                          && runs[t + 1].zB == r.zB;
        cols.push_back({r.x0, r.zA, r.zB, pair ? 0 : 1});
        if (pair) t++;
    }
    if (interior == IN_PAIRS && !D2.empty()) {
        /* segment-major, x fastest: the four items of a workgroup and the
         * workgroups an XCD runs together share their halo columns */
        const int ps = seg_rows > 0 ? (seg_rows + 3) / 4 * 4 : D2.z1 - D2.z0;
        for (int z = D2.z0; z < D2.z1; z += ps) {
            const int e = z + ps < D2.z1 ? z + ps : D2.z1;
            for (int x0 = D2.x0; x0 < D2.x1; x0 += kPairCols) {
                const int w = D2.x1 - x0 < kPairCols ? D2.x1 - x0 : kPairCols;
                inner.push_back({pair_pack(sim, w / 2), x0, z, e});
            }
        }
    }
    if (interior != IN_COLUMNS) return;
    int zmax = 0;
    for (const Col &c : cols) zmax = c.zB - c.zA > zmax ? c.zB - c.zA : zmax;
    for (int off = 0; off < zmax; off += seg)      /* segment-major, x fastest */
        for (const Col &c : cols) {
            const int z = c.zA + off;
            if (z >= c.zB) continue;
            const int e = z + seg < c.zB ? z + seg : c.zB;
            inner.push_back({col_pack(sim, c.half), c.x0, z, e});
        }
}

/* Reorder a tile list {sim, x0, z0, class} for kXcds XCDs.  Workgroup b runs
 * on XCD b % kXcds and the tile kernel gives XCD k the k-th share of the
 * list, so the shares should cost the same.  Every simulation's tiles are cut
 * into kXcds slabs of whole tile columns of equal cost (a strip tile counts
 * as three interior tiles of its size), slab k of every simulation goes to
 * share k, and
 * the shares are padded to equal length with entries {-1, 0, 0, 0} that the
 * kernel skips.  Inside each simulation's slab the classes with the longest
 * chains come first (small tiles before full ones; corner, z-strip, x-strip,
 * interior), each row-major.
 * The list must hold each simulation's tiles contiguously. */
constexpr int kXcds = 8;
constexpr double kStripCost = 3.0;

/* tz = rows of a full tile; a 64 x 4 tile (T_SMALL) counts 4/tz of one */
inline void balance_for_xcds(std::vector<WorkItem> &tiles, int tz)
{
    if (tiles.empty()) return;
    std::vector<std::vector<WorkItem>> part(kXcds);
    const double small = tz > 0 ? 4.0 / tz : 1.0;
    auto cost = [small](const WorkItem &t) {
        return ((t.d & 3) ? kStripCost : 1.0) * ((t.d & T_SMALL) ? small : 1.0);
    };
    size_t i = 0;
    while (i < tiles.size()) {
        size_t j = i;
        while (j < tiles.size() && tiles[j].a == tiles[i].a) j++;
        std::vector<WorkItem> sim(tiles.begin() + i, tiles.begin() + j);
        std::sort(sim.begin(), sim.end(), [](const WorkItem &a, const WorkItem &b) {
            return a.b != b.b ? a.b < b.b : a.c < b.c;
        });
        double total = 0.0;
        for (const WorkItem &t : sim) total += cost(t);
        double acc = 0.0;
        std::vector<std::vector<WorkItem>> slab(kXcds);
        size_t c = 0;
        while (c < sim.size()) {
            size_t e = c;
            double cw = 0.0;
            while (e < sim.size() && sim[e].b == sim[c].b) cw += cost(sim[e++]);
            int k = (int)((acc + 0.5 * cw) / total * kXcds);
            k = k < 0 ? 0 : (k >= kXcds ? kXcds - 1 : k);
            slab[k].insert(slab[k].end(), sim.begin() + c, sim.begin() + e);
            acc += cw;
            c = e;
        }
        for (int k = 0; k < kXcds; k++) {
            std::sort(slab[k].begin(), slab[k].end(), [](const WorkItem &a, const WorkItem &b) {
                if (a.d != b.d) return a.d > b.d;
                return a.c != b.c ? a.c < b.c : a.b < b.b;
            });
            part[k].insert(part[k].end(), slab[k].begin(), slab[k].end());
        }
        i = j;
    }
    size_t per = 0;
    for (const auto &p : part) per = p.size() > per ? p.size() : per;
    tiles.clear();
    for (const auto &p : part) {
        tiles.insert(tiles.end(), p.begin(), p.end());
        tiles.insert(tiles.end(), per - p.size(), WorkItem{-1, 0, 0, 0});
    }
}

}  // namespace rtm

#endif /* RTM_WORK_H */
