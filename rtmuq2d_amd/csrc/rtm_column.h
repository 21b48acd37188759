/*
 * rtm_column.h -- the interior column kernel body: one wavefront marches
 * down a strip of 128 (or 64) columns of the region that no CPML strip
 * touches, with no LDS and no barrier.
 *
 * Lane l owns the two adjacent columns x0+2l and x0+2l+1 of every row.
 * z-neighbours come from a register window of rows k-4 .. k+7 that slides
 * down four rows at a time while the next four rows are already in flight;
 * x-neighbours come from lanes l-1, l-2 (x-4 .. x-1) and l+1, l+2 (x+2 ..
 * x+5) through DPP wave shifts.  The halo columns x0-4 .. x0-1 (lane 0) and
 * x0+128 .. x0+131 (lane 63) arrive as the "old" operand of the shift, from
 * one 4-wide load per row.  In a 64-column strip lanes 32 .. 63 load the
 * next 64 columns (in bounds: the strip ends at least 4 points before the
 * right CPML) and only supply halo values.
 *
 * Per point the work is the stencil of rtm::stencil_val (identical
 * association order, so identical bits) plus, on the rows where they apply,
 * the source add, receiver replacement or recording, snapshot store and
 * imaging of rtm_tile.h's tile_points.  The CPML strips themselves are
 * stepped by the tile kernel in the same launch sequence.
 */
#ifndef RTM_COLUMN_H
#define RTM_COLUMN_H

#include "rtm_tile.h"

namespace rtm {

template <typename R>
struct alignas(2 * sizeof(R)) Pair {
    R a, b;
};

template <typename R>
struct alignas(4 * sizeof(R)) Quad {
    R x, y, z, w;
};

/* Column-strip work item {sim | half << 24, x0, zA, zB}: columns x0 ..
 * x0+127 (x0+63 when `half`), rows zA .. zB-1 (a multiple of 4 rows). */
RTM_HD int col_sim(int packed) { return packed & 0xffffff; }
RTM_HD int col_half(int packed) { return packed >> 24; }
RTM_HD int col_pack(int sim, int half) { return sim | (half << 24); }

/* The two updated points of one lane at row k.  w[0..8] are rows k-4 ..
 * k+4 of the lane's columns (x, x+1); L1 = (x-2, x-1), L2 = (x-4, x-3),
 * R1 = (x+2, x+3), R2 = (x+4, x+5). */
template <typename R>
RTM_HD Pair<R> col_update(const Pair<R> *w, Pair<R> L1, Pair<R> L2, Pair<R> R1, Pair<R> R2, Pair<R> Cp,
                          Pair<R> Op)
{
    const Pair<R> c = w[4];
    Pair<R> U;
    U.a = stencil_val(Cp.a, Op.a, c.a, L1.b, c.b, w[3].a, w[5].a, L1.a, R1.a, w[2].a, w[6].a,
                      L2.b, R1.b, w[1].a, w[7].a, L2.a, R2.a, w[0].a, w[8].a);
    U.b = stencil_val(Cp.b, Op.b, c.b, c.a, R1.a, w[3].b, w[5].b, L1.b, R1.b, w[2].b, w[6].b,
                      L1.a, R2.a, w[1].b, w[7].b, L2.b, R2.b, w[0].b, w[8].b);
    return U;
}

/* A lane's receivers for a column item, loaded before the march: the
 * receiver at each of its two columns on the receiver row (-1: none) and,
 * backward, the trace samples that replace u there.  `on_item`: the
 * receiver row is one of the item's rows (uniform). */
template <typename R>
struct ColRec {
    int ra, rb;
    R va, vb;
};

template <typename R>
RTM_HD ColRec<R> col_rec(const SimDesc<R> &S, const StepView<R> &V, int xl, bool ok, bool on_item)
{
    ColRec<R> r{-1, -1, R(0), R(0)};
    if (!on_item || !(!V.forward || V.rec_row_out)) return r;   /* uniform */
    const int xc = clampi(xl, 0, S.X - 2);
    const int ra = ldg(S.recmap, xc), rb = ldg(S.recmap, xc + 1);
    r.ra = ok ? ra : -1;
    r.rb = ok ? rb : -1;
    if (!V.forward) {
        const R va = R(ldg(V.seis_row, r.ra > 0 ? r.ra : 0)), vb = R(ldg(V.seis_row, r.rb > 0 ? r.rb : 0));
        r.va = va;
        r.vb = vb;
    }
    return r;
}

/* Source add, receiver replacement or recording, snapshot store (FP32, or
 * the 16-bit store's max |U| and quantization of u^n) and imaging for the
 * lane's two points at row k (in the order of tile_points); the row tests
 * are uniform across the wavefront.  uc = u^n at the two points.  Only the
 * imaging loads memory, and only the IMG instance images (the item's step
 * is an imaging step): a load in a branch that some rows take would make
 * the march wait for all its prefetched rows at every row after it. */
template <typename R, bool Q16 = false, bool IMG = true>
RTM_HD void col_extras(const SimDesc<R> &S, const StepView<R> &V, const ColRec<R> &rec, int k, int xl, bool ok,
                       int src_row, Pair<R> &U, Pair<R> Cp, Pair<R> uc, float &umax)
{
    if (V.forward) {
        if (k == src_row) {
            const int j = k * S.P + xl;
            if (j == S.src_j) U.a = U.a + V.src_w * Cp.a;
            if (j + 1 == S.src_j) U.b = U.b + V.src_w * Cp.b;
        }
    }
    if (k == S.rec_row && (!V.forward || V.rec_row_out)) {
        if (V.forward) {
            if (rec.ra >= 0) stg(V.rec_row_out, rec.ra, float(U.a));
            if (rec.rb >= 0) stg(V.rec_row_out, rec.rb, float(U.b));
        } else {
            U.a = rec.ra >= 0 ? rec.va : U.a;
            U.b = rec.rb >= 0 ? rec.vb : U.b;
        }
    }
    const int kk = k - S.TU;
    if (!(Q16 ? (V.qsnap || V.smax) : (V.snap || V.sfield)) || kk < 0 || kk >= S.nz) return;
    if (!ok) return;
    const R vals[2] = {U.a, U.b};
    const R us[2] = {uc.a, uc.b};
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int ii = xl + e - S.TL;
        if (ii < 0 || ii >= S.nx) continue;
        const int q = kk * S.nx + ii;
        if (V.forward) {
            if (!Q16) {
                if (V.snap) stg_stream(V.snap, q, vals[e]);
            } else {
                if (V.smax) umax = fmaxf(umax, fabsf(float(vals[e])));
                if (V.qsnap) stg_stream(V.qsnap, q, quant16(float(us[e]), V.qscale, S.qmode));
            }
        } else if (IMG) {
            const R s = Q16        ? R(dequant16(ldg(V.qsnap, q), V.qscale, S.qmode))
                        : V.sfield ? ldg(V.sfield, k * S.P + xl + e)
                                   : ldg(V.snap, q);
            if (S.flags & F_KAHAN) {
                R I = ldg(S.I, q), Ic = ldg(S.Ic, q), Is = ldg(S.Is, q), Isc = ldg(S.Isc, q);
                kahan_add(I, Ic, s * vals[e]);
                kahan_add(Is, Isc, s * s);
                stg(S.I, q, I); stg(S.Ic, q, Ic); stg(S.Is, q, Is); stg(S.Isc, q, Isc);
            } else {
                stg(S.I, q, ldg(S.I, q) + s * vals[e]);
                stg(S.Is, q, ldg(S.Is, q) + s * s);
            }
        }
    }
}

/* Source row of a sim this step (-1 when no source is added). */
template <typename R>
RTM_HD int col_src_row(const SimDesc<R> &S, const StepView<R> &V)
{
    return (V.forward && V.src) ? S.src_j / S.P : -1;
}

}  // namespace rtm

#endif /* RTM_COLUMN_H */
