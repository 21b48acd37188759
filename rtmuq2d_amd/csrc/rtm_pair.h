/*
 * rtm_pair.h -- two time steps per launch over the inner part of the image
 * region (register-blocked, no LDS, no barrier).
 *
 * The region D2 is a rectangle of whole tiles inside the image region
 * [TL, X-TR) x [TU, Z-TD).  Every point within 4 of D2 then lies outside
 * the CPML strips, so u^{n+1} there is the plain leapfrog update of
 * SRC/stencilDFM.c:18-28 plus the source add or receiver replacement
 * (SRC/forward_modeling.c:105-109, SRC/backward_modeling.c:106-110): one
 * wavefront can compute u^{n+1} on D2 widened by 4 and then u^{n+2} on D2
 * from registers, reading u^n, u^{n-1} and C once for two steps.
 *
 * A pair item {sim | wp << 24, x0, zA, zB} covers output columns
 * x0 .. x0+2wp-1 (wp <= 56) and rows zA .. zB-1 of D2.  Lane l holds
 * columns x0-8+2l and x0-7+2l (128 columns); u^{n+1} is right on lanes
 * 2..61 (its x-neighbours are two lanes away) and u^{n+2} on lanes 4..59,
 * the item's columns.  The wavefront marches down in chunks of four rows:
 * u^{n+1} on rows m0 .. m0+3 from the u^n window, then u^{n+2} on rows
 * m0-4 .. m0-1 from the u^{n+1} window.
 *
 * The tile kernel steps the rest of the grid one step per launch, so a
 * launch pair is {pair items at n and n+1, tiles at n} then {tiles at n+1}.
 * u lives in four buffers (u^n in ubuf[n & 3]): each launch reads two time
 * levels and writes others, so no launch overwrites what another item of
 * the same launch still reads.
 *
 * Per point the arithmetic is rtm::stencil_val and the extras of
 * rtm_tile.h's tile_points (identical association order, identical bits).
 */
#ifndef RTM_PAIR_H
#define RTM_PAIR_H

#include "rtm_column.h"

namespace rtm {

constexpr int kPairCols = 112;   /* output columns of a full pair item */

RTM_HD int pair_sim(int packed) { return packed & 0xffffff; }
RTM_HD int pair_wp(int packed) { return packed >> 24; }          /* output column pairs */
RTM_HD int pair_pack(int sim, int wp) { return sim | (wp << 24); }

/* D2 for a grid whose tiles are 64 x TZ from (4, 4): the largest rectangle
 * of whole tiles inside the image region with u^n readable 8 points around
 * it and u^{n+1} computed 4 points around it inside the stepped region.
 * Empty (x1 <= x0 or z1 <= z0) when the grid has no such tiles. */
struct PairRegion {
    int x0, x1, z0, z1;
    bool empty() const { return x1 <= x0 || z1 <= z0; }
};

inline PairRegion pair_region(int X, int Z, int TL, int TR, int TU, int TD, int TZ)
{
    auto up = [](int v, int o, int s) { return v <= o ? o : o + (v - o + s - 1) / s * s; };
    auto down = [](int v, int o, int s) { return v < o ? o - s : o + (v - o) / s * s; };
    PairRegion r;
    r.x0 = up(TL > 2 * kHalf ? TL : 2 * kHalf, kHalf, 64);
    r.x1 = down(X - TR < X - 2 * kHalf ? X - TR : X - 2 * kHalf, kHalf, 64);
    r.z0 = up(TU > 2 * kHalf ? TU : 2 * kHalf, kHalf, TZ);
    r.z1 = down(Z - TD < Z - 2 * kHalf ? Z - TD : Z - 2 * kHalf, kHalf, TZ);
    return r;
}

/* The bounds k_pairs relies on for an item of a grid with X x Z points and
 * row pitch P >= X: u^n rows zA-8 .. zB+7 and columns x0-8 .. x0+2wp+7 in
 * the grid, u^{n+1} rows zA-4 .. zB+3 and columns x0-4 .. x0+2wp+3 in the
 * stepped region, a row count that is a positive multiple of 4, even x0
 * and 1 <= wp <= 56. */
inline bool pair_item_ok(int wp, int x0, int zA, int zB, int X, int Z)
{
    return wp >= 1 && 2 * wp <= kPairCols && (x0 & 1) == 0 && x0 - 2 * kHalf >= 0 && x0 + 2 * wp + 2 * kHalf <= X &&
           zA - 2 * kHalf >= 0 && zB + 2 * kHalf <= Z && zB > zA && (zB - zA) % 4 == 0;
}

/* A lane's receiver operands for the item, loaded before the march: the
 * receiver at each of its two columns on the receiver row (-1: none) and,
 * backward, the trace samples that replace u there after each step. */
template <typename R>
struct PairRec {
    int ra, rb;
    R va1, vb1, va2, vb2;
};

template <typename R>
RTM_HD PairRec<R> pair_rec(const SimDesc<R> &S, const StepView<R> &V1, const StepView<R> &V2, bool two, int xl,
                           bool on_row)
{
    PairRec<R> r{-1, -1, R(0), R(0), R(0), R(0)};
    if (!on_row || !(!V1.forward || V1.rec_row_out)) return r;   /* uniform */
    const int xc = clampi(xl, 0, S.X - 2);
    r.ra = ldg(S.recmap, xc);
    r.rb = ldg(S.recmap, xc + 1);
    if (!V1.forward) {
        const int ia = r.ra > 0 ? r.ra : 0, ib = r.rb > 0 ? r.rb : 0;
        r.va1 = R(ldg(V1.seis_row, ia));
        r.vb1 = R(ldg(V1.seis_row, ib));
        if (two) {
            r.va2 = R(ldg(V2.seis_row, ia));
            r.vb2 = R(ldg(V2.seis_row, ib));
        }
    }
    return r;
}

/* A non-temporal store through the global address space (a flat store would
 * also count against the scalar memory counter). */
template <typename T>
RTM_HD void stg_nt(T *p, int i, T v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned off = (unsigned)i * (unsigned)sizeof(T);
    __builtin_nontemporal_store(v, (RTM_GAS T *)((RTM_GAS char *)p + off));
#else
    p[i] = v;
#endif
}

/* Source add, receiver replacement or recording, snapshot store and
 * imaging of one lane's two points (xl, xl+1) at row k after one of the two
 * steps (view V; `second` selects the step's trace samples in rec), in the
 * order of tile_points.  The replacement and the source add apply on every
 * lane, since neighbouring lanes read these values; recording, snapshot and
 * imaging only on the item's own points (`own`).  Only the imaging loads
 * memory (the instance IMG); everything else the march needs was loaded
 * before it, so no load of a branch taken on some rows only makes the
 * march wait for its prefetched rows. */
template <typename R, bool IMG>
RTM_HD void pair_extras(const SimDesc<R> &S, const StepView<R> &V, const PairRec<R> &rec, bool second, int k, int xl,
                        bool own, int srow, Pair<R> &U, Pair<R> Cp)
{
    if (V.forward && k == srow) {
        const int j = k * S.P + xl;
        if (j == S.src_j) U.a = U.a + V.src_w * Cp.a;
        if (j + 1 == S.src_j) U.b = U.b + V.src_w * Cp.b;
    }
    if (k == S.rec_row) {
        if (V.forward) {
            if (own && rec.ra >= 0) stg(V.rec_row_out, rec.ra, float(U.a));
            if (own && rec.rb >= 0) stg(V.rec_row_out, rec.rb, float(U.b));
        } else {
            U.a = rec.ra >= 0 ? (second ? rec.va2 : rec.va1) : U.a;
            U.b = rec.rb >= 0 ? (second ? rec.vb2 : rec.vb1) : U.b;
        }
    }
    if (!own || !V.snap) return;
    const int q = (k - S.TU) * S.nx + (xl - S.TL);   /* own points lie in the image region */
    const R vals[2] = {U.a, U.b};
#pragma unroll
    for (int e = 0; e < 2; e++) {
        if (V.forward) {
            stg_nt(V.snap, q + e, vals[e]);
        } else if (IMG) {
            const R s = ldg(V.snap, q + e);
            if (S.flags & F_KAHAN) {
                R I = ldg(S.I, q + e), Ic = ldg(S.Ic, q + e), Is = ldg(S.Is, q + e), Isc = ldg(S.Isc, q + e);
                kahan_add(I, Ic, s * vals[e]);
                kahan_add(Is, Isc, s * s);
                stg(S.I, q + e, I); stg(S.Ic, q + e, Ic); stg(S.Is, q + e, Is); stg(S.Isc, q + e, Isc);
            } else {
                stg(S.I, q + e, ldg(S.I, q + e) + s * vals[e]);
                stg(S.Is, q + e, ldg(S.Is, q + e) + s * s);
            }
        }
    }
}

}  // namespace rtm

#endif /* RTM_PAIR_H */
