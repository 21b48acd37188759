/*
 * rtm_work.h -- how one simulation's stepped region 4 <= i < X-4,
 * 4 <= k < Z-4 is split into work items for the step kernels.  Shared by
 * the library and the CPU test harness so both walk the same items.
 *
 * Tiles are 64 x TZ (TZ = 4*rz) from (4, 4).  Each tile has a class: which
 * CPML strips its points touch.  The interior (class-0 tiles, whose points
 * need no CPML term) is handled by one of two kernels:
 *   IN_TILES    every tile goes to the tile kernel;
 *   IN_COLUMNS  class-0 runs of two adjacent tile columns with the same
 *               rows become one 128-column item, a lone column a 64-column
 *               item, in segments of at most seg_rows rows (rtm_column.h).
 */
#ifndef RTM_WORK_H
#define RTM_WORK_H

#include <vector>

#include "rtm_column.h"

namespace rtm {

enum Interior : int { IN_TILES = 0, IN_COLUMNS = 1 };

struct WorkItem {
    int a, b, c, d;   /* tile {sim, x0, z0, class}; interior {sim(+half), x0, zA, zB} */
};

inline void build_work(std::vector<WorkItem> &tiles, std::vector<WorkItem> &inner, int sim, int X, int Z, int TL,
                       int TR, int TU, int TD, int rz, int interior, int seg_rows)
{
    const int TZ = 4 * rz;
    SimDesc<float> d{};
    d.X = X; d.Z = Z; d.TL = TL; d.TR = TR; d.TU = TU; d.TD = TD;
    int seg = 0;
    if (interior == IN_COLUMNS) seg = seg_rows > 0 ? (seg_rows + 3) / 4 * 4 : 1 << 30;
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
            tiles.push_back({sim, x0, z0, tile_class(d, x0, z0, TZ)});
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

}  // namespace rtm

#endif /* RTM_WORK_H */
