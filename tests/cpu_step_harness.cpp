// cpu_step_harness.cpp -- runs the device step arithmetic (rtm_step.h) on the
// CPU so the per-point fused update can be checked against the oracle
// without a GPU.  Built by tests/test_cpu_harness.py with
// g++ -O2 -ffp-contract=off; test infrastructure only.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <vector>

#include "rtm_step.h"
#include "rtm_tile.h"
#include "rtm_work.h"

// interior tiles run through the clamped [0] and the unclamped (SAFE) [1]
// phases since the last hx_interior_tiles call
static long g_cls0_tiles[2] = {0, 0};

template <typename R>
static void run(const int *g, const R *hs, const R *u, R *U, const R *Cf, const R *ax, const R *bx,
                const R *az, const R *bz, const R *pXi, R *pXo, R *cX, const R *pZi, R *pZo, R *cZ,
                const int *recmap, const float *seis_row, R *snap, R *I, R *Is, R *Ic, R *Isc)
{
    rtm::StepArgs<R> A{};
    A.X = g[0]; A.Z = g[1]; A.P = g[2];
    A.TL = g[3]; A.TR = g[4]; A.TU = g[5]; A.TD = g[6];
    A.nx = g[7]; A.nz = g[8]; A.mode = g[9]; A.src_j = g[10]; A.rec_row = g[11]; A.nrec = g[12];
    A.h1 = hs[0]; A.h2 = hs[1]; A.h3 = hs[2]; A.src_w = hs[3];
    A.u = u; A.U = U; A.C = Cf;
    A.ax = ax; A.bx = bx; A.az = az; A.bz = bz;
    A.psiXi = pXi; A.psiXo = pXo; A.csiX = cX;
    A.psiZi = pZi; A.psiZo = pZo; A.csiZ = cZ;
    A.recmap = recmap; A.seis_row = seis_row;
    A.snap = snap; A.I = I; A.Is = Is; A.Ic = Ic; A.Isc = Isc;
    for (int k = rtm::kHalf; k < A.Z - rtm::kHalf; k++)
        for (int i = rtm::kHalf; i < A.X - rtm::kHalf; i++) rtm::step_point(A, i, k);
    if (g[13])   // outer-column top-strip pass, as k_top_edges
        for (int k = rtm::kHalf; k < A.TU - rtm::kHalf; k++)
            for (int c = 0; c < 8; c++) rtm::cpml_top_edge_point(A, c < 4 ? c : A.X - 8 + c, k);
}

// The kernels' max |U| commit (a wavefront max and an atomic max on the
// float bits) done sequentially.
template <typename R>
static void commit_max(const rtm::StepView<R> &V, float m)
{
    if (V.smax && rtm::f32_bits(m) > *V.smax) *V.smax = rtm::f32_bits(m);
}

// Emulation of k_tiles for one simulation at step n: every tile, phase by
// phase, all 64*NW threads of a phase before the next (the barriers).
template <typename R, int RZ, int CLS, bool Q16, bool I16 = false, int NW = 4>
static void run_tile(rtm::TileLds<R, RZ, NW> &L, rtm::TileRegs<R, RZ, NW> *G, const rtm::SimDesc<R> &S,
                     const rtm::StepView<R> &V, int x0, int z0);

// An interior tile runs the unclamped (SAFE) phases when rtm::tile_safe
// holds, as k_tiles dispatches it.
template <typename R, int RZ, int CLS, bool Q16, bool I16, int NW, bool SAFE>
static void run_tile_s(rtm::TileLds<R, RZ, NW> &L, rtm::TileRegs<R, RZ, NW> *G, const rtm::SimDesc<R> &S,
                       const rtm::StepView<R> &V, int x0, int z0)
{
    // forward steps as the store instance steps them (no imaging operands,
    // no imaging code), backward steps as the full one
    for (int ty = 0; ty < NW; ty++)
        for (int tx = 0; tx < 64; tx++) {
            if (V.forward) rtm::tile_fetch<R, RZ, CLS, Q16, I16, NW, false, SAFE>(G[ty * 64 + tx], S, V, x0, z0, tx, ty);
            else rtm::tile_fetch<R, RZ, CLS, Q16, I16, NW, true, SAFE>(G[ty * 64 + tx], S, V, x0, z0, tx, ty);
            rtm::tile_stage<R, RZ, CLS, NW>(L, G[ty * 64 + tx], tx, ty);
        }
    if (CLS)
        for (int ty = 0; ty < NW; ty++)
            for (int tx = 0; tx < 64; tx++)
                rtm::tile_psi<R, RZ, CLS, NW>(L, G[ty * 64 + tx], S, x0, z0, tx, ty);
    float m = 0.0f;
    for (int ty = 0; ty < NW; ty++)
        for (int tx = 0; tx < 64; tx++) {
            const float t =
                V.forward
                    ? rtm::tile_points<R, RZ, CLS, Q16, I16, NW, 1, SAFE>(L, G[ty * 64 + tx], S, V, x0, z0, tx, ty)
                    : rtm::tile_points<R, RZ, CLS, Q16, I16, NW, 2, SAFE>(L, G[ty * 64 + tx], S, V, x0, z0, tx, ty);
            m = t > m ? t : m;
        }
    commit_max(V, m);
    if (I16 && V.forward && V.qsnap) {
        // the kernel's wavefront max of each row's |U| (the block), then the store
        for (int ty = 0; ty < NW; ty++) {
            float rmax[RZ];
            for (int rr = 0; rr < RZ; rr++) {
                rmax[rr] = 0.0f;
                for (int tx = 0; tx < 64; tx++) {
                    const float a = std::fabs(float(G[ty * 64 + tx].uq[rr]));
                    rmax[rr] = a > rmax[rr] ? a : rmax[rr];
                }
            }
            for (int tx = 0; tx < 64; tx++)
                rtm::tile_store_i16<R, RZ, NW>(G[ty * 64 + tx], S, V, rmax, x0, z0, tx, ty);
        }
    }
}

template <typename R, int RZ, int CLS, bool Q16, bool I16, int NW>
static void run_tile(rtm::TileLds<R, RZ, NW> &L, rtm::TileRegs<R, RZ, NW> *G, const rtm::SimDesc<R> &S,
                     const rtm::StepView<R> &V, int x0, int z0)
{
    if (CLS == 0 && rtm::tile_safe(S, x0, z0, NW * RZ)) {
        g_cls0_tiles[1]++;
        run_tile_s<R, RZ, 0, Q16, I16, NW, true>(L, G, S, V, x0, z0);
    } else {
        if (CLS == 0) g_cls0_tiles[0]++;
        run_tile_s<R, RZ, CLS, Q16, I16, NW, false>(L, G, S, V, x0, z0);
    }
}

// Emulation of k_columns for one item: every row, all 64 lanes, with the
// DPP lane shifts done as array shifts (lane 0 / 63 take the halo quad).
template <typename R, bool Q16>
static void run_column_item(const rtm::SimDesc<R> &S, const rtm::StepView<R> &V, int packed, int x0, int zA, int zB)
{
    using P2 = rtm::Pair<R>;
    const int P = S.P;
    const int half = rtm::col_half(packed);
    const int srow = rtm::col_src_row(S, V);
    float umax = 0.0f;
    auto at = [&](const R *f, int k, int x) { return f[(size_t)k * P + x]; };
    const bool on_item = S.rec_row >= zA && S.rec_row < zB;
    for (int k = zA; k < zB; k++) {
        P2 c[64], L1[64], L2[64], R1[64], R2[64];
        rtm::Quad<R> H[64];
        int xl[64], xc[64];
        for (int l = 0; l < 64; l++) {
            xl[l] = x0 + 2 * l;
            xc[l] = xl[l] < P - 2 ? xl[l] : P - 2;
            c[l] = {at(V.u, k, xc[l]), at(V.u, k, xc[l] + 1)};
            const int hx = rtm::clampi(l == 63 ? x0 + 128 : x0 - 4, 0, P - 4);
            H[l] = {at(V.u, k, hx), at(V.u, k, hx + 1), at(V.u, k, hx + 2), at(V.u, k, hx + 3)};
        }
        for (int l = 0; l < 64; l++) L1[l] = l > 0 ? c[l - 1] : P2{H[0].z, H[0].w};
        for (int l = 0; l < 64; l++) L2[l] = l > 0 ? L1[l - 1] : P2{H[0].x, H[0].y};
        for (int l = 0; l < 64; l++) R1[l] = l < 63 ? c[l + 1] : P2{H[63].x, H[63].y};
        for (int l = 0; l < 64; l++) R2[l] = l < 63 ? R1[l + 1] : P2{H[63].z, H[63].w};
        for (int l = 0; l < 64; l++) {
            P2 w[9];
            for (int r = 0; r < 9; r++) w[r] = {at(V.u, k - 4 + r, xc[l]), at(V.u, k - 4 + r, xc[l] + 1)};
            const P2 Cp = {at(S.C, k, xc[l]), at(S.C, k, xc[l] + 1)};
            const P2 Op = {at(V.U, k, xc[l]), at(V.U, k, xc[l] + 1)};
            P2 U = rtm::col_update(w, L1[l], L2[l], R1[l], R2[l], Cp, Op);
            const bool ok = l < (half ? 32 : 64);
            rtm::col_extras<R, Q16>(S, V, rtm::col_rec(S, V, xl[l], ok, on_item), k, xl[l], ok, srow, U, Cp, w[4],
                                    umax);
            if (ok) {
                V.U[(size_t)k * P + xl[l]] = U.a;
                V.U[(size_t)k * P + xl[l] + 1] = U.b;
            }
        }
    }
    commit_max(V, umax);
}

// Emulation of k_pairs for one item {sim | wp << 24, x0, zA, zB}: u^{n+1}
// on rows zA-4 .. zB+3 for all 64 lanes, then u^{n+2} on rows zA .. zB-1,
// with the DPP lane shifts done as array shifts (a lane with no neighbour
// keeps its own value, as the kernel's shifts do).  `two`: the second step
// runs (it is part of the launch and the simulation is still active).
template <typename R>
static void run_pair_item(const rtm::SimDesc<R> &S, const rtm::StepView<R> &V1, const rtm::StepView<R> &V2,
                          bool two, int packed, int x0, int zA, int zB)
{
    using P2 = rtm::Pair<R>;
    const int P = S.P, wp = rtm::pair_wp(packed);
    const int srow1 = rtm::col_src_row(S, V1), srow2 = two ? rtm::col_src_row(S, V2) : -1;
    const bool on_row = S.rec_row >= zA - 4 && S.rec_row < zB + 4;
    const int m_lo = zA - 4, m_hi = zB + 4;
    std::vector<P2> U1((size_t)(m_hi - m_lo) * 64);
    auto row1 = [&](int m) { return &U1[(size_t)(m - m_lo) * 64]; };
    int xl[64], xc[64];
    for (int l = 0; l < 64; l++) {
        xl[l] = x0 - 8 + 2 * l;
        xc[l] = rtm::clampi(xl[l], 0, P - 2);
    }
    auto shifts = [](const P2 *c, P2 *L1, P2 *L2, P2 *R1, P2 *R2) {
        for (int l = 0; l < 64; l++) L1[l] = l > 0 ? c[l - 1] : c[l];
        for (int l = 0; l < 64; l++) L2[l] = l > 0 ? L1[l - 1] : L1[l];
        for (int l = 0; l < 64; l++) R1[l] = l < 63 ? c[l + 1] : c[l];
        for (int l = 0; l < 64; l++) R2[l] = l < 63 ? R1[l + 1] : R1[l];
    };
    auto at = [&](const R *f, int k, int x) { return P2{f[(size_t)k * P + x], f[(size_t)k * P + x + 1]}; };
    for (int m = m_lo; m < m_hi; m++) {
        P2 c[64], L1[64], L2[64], R1[64], R2[64];
        for (int l = 0; l < 64; l++) c[l] = at(V1.u, m, xc[l]);
        shifts(c, L1, L2, R1, R2);
        for (int l = 0; l < 64; l++) {
            P2 w[9];
            for (int r = 0; r < 9; r++) w[r] = at(V1.u, m - 4 + r, xc[l]);
            const P2 Cp = at(S.C, m, xc[l]);
            P2 U = rtm::col_update(w, L1[l], L2[l], R1[l], R2[l], Cp, at(V1.Uo, m, xc[l]));
            const bool own = l >= 4 && l < 4 + wp && m >= zA && m < zB;
            rtm::pair_extras<R, true>(S, V1, rtm::pair_rec(S, V1, V2, two, xl[l], on_row), false, m, xl[l], own,
                                      srow1, U, Cp);
            row1(m)[l] = U;
            if (own) {
                V1.U[(size_t)m * P + xl[l]] = U.a;
                V1.U[(size_t)m * P + xl[l] + 1] = U.b;
            }
        }
    }
    if (!two) return;
    for (int q = zA; q < zB; q++) {
        P2 L1[64], L2[64], R1[64], R2[64];
        shifts(row1(q), L1, L2, R1, R2);
        for (int l = 0; l < 64; l++) {
            P2 w[9];
            for (int r = 0; r < 9; r++) w[r] = row1(q - 4 + r)[l];
            const P2 Cp = at(S.C, q, xc[l]);
            P2 U = rtm::col_update(w, L1[l], L2[l], R1[l], R2[l], Cp, at(V1.u, q, xc[l]));
            const bool own = l >= 4 && l < 4 + wp;
            rtm::pair_extras<R, true>(S, V2, rtm::pair_rec(S, V1, V2, two, xl[l], on_row), true, q, xl[l], own,
                                      srow2, U, Cp);
            if (own) {
                V2.U[(size_t)q * P + xl[l]] = U.a;
                V2.U[(size_t)q * P + xl[l] + 1] = U.b;
            }
        }
    }
}

// One launch of one simulation over the work items of the library's builder
// (rtm::build_work): tiles, plus column interior items, or (interior
// IN_PAIRS) the pair items at steps n and n+1 when `pair_launch` (the first
// launch of a pair; `two`: its second step is part of the run).
template <typename R, int RZ, bool Q16, bool I16 = false>
static void run_tiles_q(const rtm::SimDesc<R> &S, int n, int interior, int seg_rows, int pair_launch = 0,
                        bool two = false)
{
    const rtm::StepView<R> V = rtm::step_view(S, n);
    if (!V.active) return;
    auto *L = new rtm::TileLds<R, RZ>();
    auto *G = new rtm::TileRegs<R, RZ>[256];
    auto *L8 = new rtm::TileLds<R, RZ, 8>();
    auto *G8 = new rtm::TileRegs<R, RZ, 8>[512];
    auto *Ls = new rtm::TileLds<R, 1>();
    auto *Gs = new rtm::TileRegs<R, 1>[256];
    std::vector<rtm::WorkItem> tiles, inner;
    rtm::build_work(tiles, inner, 0, S.X, S.Z, S.TL, S.TR, S.TU, S.TD, RZ, interior, seg_rows);
    for (const auto &t : tiles) {
        // the LDS holds stale data from the previous tile, as on the GPU
        if (t.d & rtm::T_SMALL) {   // a 64 x 4 tile: of a mixed list (k_tiles_mix), or a corner piece
            switch (t.d & 3) {
            case 0: run_tile<R, 1, 0, Q16, I16>(*Ls, Gs, S, V, t.b, t.c); break;
            case 1: run_tile<R, 1, 1, Q16, I16>(*Ls, Gs, S, V, t.b, t.c); break;
            case 2: run_tile<R, 1, 2, Q16, I16>(*Ls, Gs, S, V, t.b, t.c); break;
            default: run_tile<R, 1, 3, Q16, I16>(*Ls, Gs, S, V, t.b, t.c); break;
            }
            continue;
        }
        if (interior == rtm::IN_TILES8) {   // workgroups of 8 wavefronts (k_tiles NW = 8)
            switch (t.d) {
            case 0: run_tile<R, RZ, 0, Q16, I16, 8>(*L8, G8, S, V, t.b, t.c); break;
            case 1: run_tile<R, RZ, 1, Q16, I16, 8>(*L8, G8, S, V, t.b, t.c); break;
            case 2: run_tile<R, RZ, 2, Q16, I16, 8>(*L8, G8, S, V, t.b, t.c); break;
            default: run_tile<R, RZ, 3, Q16, I16, 8>(*L8, G8, S, V, t.b, t.c); break;
            }
            continue;
        }
        switch (t.d) {
        case 0: run_tile<R, RZ, 0, Q16, I16>(*L, G, S, V, t.b, t.c); break;
        case 1: run_tile<R, RZ, 1, Q16, I16>(*L, G, S, V, t.b, t.c); break;
        case 2: run_tile<R, RZ, 2, Q16, I16>(*L, G, S, V, t.b, t.c); break;
        default:
            if constexpr (rtm::corner_tiles_small(RZ, 4)) {
                // the kernels' LDS holds one psi plane here; the lists cut corners
                fprintf(stderr, "harness: a %d-row corner tile in the list\n", 4 * RZ);
                abort();
            } else {
                run_tile<R, RZ, 3, Q16, I16>(*L, G, S, V, t.b, t.c);
            }
            break;
        }
    }
    if (interior == rtm::IN_PAIRS) {
        if (pair_launch) {
            const rtm::StepView<R> V2 = rtm::step_view(S, n + 1);
            for (const auto &it : inner) run_pair_item<R>(S, V, V2, two && V2.active, it.a, it.b, it.c, it.d);
        }
    } else {
        for (const auto &it : inner) run_column_item<R, Q16>(S, V, it.a, it.b, it.c, it.d);
    }
    delete[] Gs;
    delete Ls;
    delete[] G8;
    delete L8;
    delete[] G;
    delete L;
}

template <typename R, int RZ>
static void run_tiles(const rtm::SimDesc<R> &S, int n, int interior, int seg_rows, int pair_launch, bool two)
{
    if (S.qmode == rtm::Q_I16) run_tiles_q<R, RZ, false, true>(S, n, interior, seg_rows);
    else if (S.qmode) run_tiles_q<R, RZ, true>(S, n, interior, seg_rows);
    else run_tiles_q<R, RZ, false>(S, n, interior, seg_rows, pair_launch, two);
}

template <typename R>
static void tiles_entry(const int *g, const R *hs, R *const *ptr, const int *recmap, const float *wav,
                        const float *seis_t, float *seis_rec)
{
    rtm::SimDesc<R> S{};
    S.X = g[0]; S.Z = g[1]; S.P = g[2];
    S.TL = g[3]; S.TR = g[4]; S.TU = g[5]; S.TD = g[6];
    S.nx = g[7]; S.nz = g[8]; S.nt = g[9]; S.TS = g[10]; S.nw = g[11]; S.nrec = g[12];
    S.phase = g[13]; S.flags = g[14]; S.src_j = g[15]; S.rec_row = g[16];
    const int n = g[17], rz = g[18], seg_rows = g[19], interior = g[20];
    S.qmode = g[21];
    S.q16 = (unsigned short *)(void *)ptr[18];
    S.smax = (unsigned *)(void *)ptr[19];
    if (S.qmode == rtm::Q_I16) {   // the int16 block store: its exponents in the max-slot entry
        S.smax = nullptr;
        S.qexp = (signed char *)(void *)ptr[19];
        S.qb0 = rtm::i16_first_block(S.TL);
        S.nqb = rtm::i16_blocks(S.TL, S.nx);
    }
    S.h1 = hs[0]; S.h2 = hs[1]; S.h3 = hs[2];
    S.ubuf[0] = ptr[0]; S.ubuf[1] = ptr[1]; S.C = ptr[2];
    S.umask = 1;
    // two-step launches (IN_PAIRS): four u buffers; g[22] = 1 for the first
    // launch of a pair, g[23] = 1 when its second step is part of the run
    const int pair_launch = interior == rtm::IN_PAIRS ? g[22] : 0;
    const bool two = interior == rtm::IN_PAIRS && g[23];
    if (interior == rtm::IN_PAIRS) {
        S.ubuf[2] = ptr[20];
        S.ubuf[3] = ptr[21];
        S.umask = 3;
    }
    S.ax = ptr[3]; S.bx = ptr[4]; S.az = ptr[5]; S.bz = ptr[6];
    S.psiX[0] = ptr[7]; S.psiX[1] = ptr[8]; S.psiZ[0] = ptr[9]; S.psiZ[1] = ptr[10];
    S.csiX = ptr[11]; S.csiZ = ptr[12];
    S.snaps = ptr[13]; S.I = ptr[14]; S.Is = ptr[15]; S.Ic = ptr[16]; S.Isc = ptr[17];
    S.recmap = recmap; S.wav = wav; S.seis_t = seis_t; S.seis_rec = seis_rec;
    switch (rz) {
    case 1: run_tiles<R, 1>(S, n, interior, seg_rows, pair_launch, two); break;
    case 2: run_tiles<R, 2>(S, n, interior, seg_rows, pair_launch, two); break;
    case 8: run_tiles<R, 8>(S, n, interior, seg_rows, pair_launch, two); break;
    default: run_tiles<R, 4>(S, n, interior, seg_rows, pair_launch, two); break;
    }
}

extern "C" {
/* interior tiles emulated through the clamped (out[0]) and the SAFE
 * (out[1]) phases since the last call */
void hx_interior_tiles(long *out)
{
    out[0] = g_cls0_tiles[0];
    out[1] = g_cls0_tiles[1];
    g_cls0_tiles[0] = g_cls0_tiles[1] = 0;
}

/* rtm::build_work's tile list for one grid as {sim, x0, z0, class}
 * quadruples (capacity `cap` items); returns its length or -1. */
int hx_work_list(int X, int Z, int TL, int TR, int TU, int TD, int rz, int interior, int seg_rows, int *out, int cap)
{
    std::vector<rtm::WorkItem> tiles, inner;
    rtm::build_work(tiles, inner, 0, X, Z, TL, TR, TU, TD, rz, interior, seg_rows);
    if ((int)tiles.size() > cap) return -1;
    for (size_t i = 0; i < tiles.size(); i++) {
        out[4 * i] = tiles[i].a; out[4 * i + 1] = tiles[i].b; out[4 * i + 2] = tiles[i].c; out[4 * i + 3] = tiles[i].d;
    }
    return (int)tiles.size();
}

/* rtm::build_work's interior items (column or pair items) for one grid as
 * {packed sim, x0, zA, zB} quadruples (capacity `cap` items); returns its
 * length or -1. */
int hx_inner_list(int X, int Z, int TL, int TR, int TU, int TD, int rz, int interior, int seg_rows, int *out, int cap)
{
    std::vector<rtm::WorkItem> tiles, inner;
    rtm::build_work(tiles, inner, 0, X, Z, TL, TR, TU, TD, rz, interior, seg_rows);
    if ((int)inner.size() > cap) return -1;
    for (size_t i = 0; i < inner.size(); i++) {
        out[4 * i] = inner[i].a; out[4 * i + 1] = inner[i].b; out[4 * i + 2] = inner[i].c; out[4 * i + 3] = inner[i].d;
    }
    return (int)inner.size();
}

/* rtm::balance_for_xcds over the tile lists of nsims same-sized grids (sim
 * ids 0..nsims-1, tile interior mode): out receives the reordered list as
 * {sim, x0, z0, class} quadruples (capacity `cap` items); returns its length
 * or -1. */
int hx_balance(int X, int Z, int TL, int TR, int TU, int TD, int rz, int nsims, int interior, int *out, int cap)
{
    std::vector<rtm::WorkItem> tiles, inner;
    for (int s = 0; s < nsims; s++) rtm::build_work(tiles, inner, s, X, Z, TL, TR, TU, TD, rz, interior, 0);
    rtm::balance_for_xcds(tiles, rtm::tile_rows(rz, interior));
    if ((int)tiles.size() > cap) return -1;
    for (size_t i = 0; i < tiles.size(); i++) {
        out[4 * i] = tiles[i].a; out[4 * i + 1] = tiles[i].b; out[4 * i + 2] = tiles[i].c; out[4 * i + 3] = tiles[i].d;
    }
    return (int)tiles.size();
}

/* the library's final-snapshot fill of the 16-bit store (k_quantize_interior) */
void hx_quantize_interior(const float *u, int P, int TL, int TU, int nx, int nz, unsigned short *q,
                          const unsigned *slots, int mode)
{
    const float scale = rtm::pow2f(14 - rtm::snap_exp(rtm::snap_max_bits(slots)));
    for (int kk = 0; kk < nz; kk++)
        for (int ii = 0; ii < nx; ii++)
            q[(size_t)kk * nx + ii] = rtm::quant16(u[(size_t)(kk + TU) * P + ii + TL], scale, mode);
}

void hx_tiles_f32(const int *g, const float *hs, float *const *ptr, const int *recmap, const float *wav,
                  const float *seis_t, float *seis_rec)
{
    tiles_entry<float>(g, hs, ptr, recmap, wav, seis_t, seis_rec);
}

void hx_tiles_f64(const int *g, const double *hs, double *const *ptr, const int *recmap,
                  const float *wav, const float *seis_t, float *seis_rec)
{
    tiles_entry<double>(g, hs, ptr, recmap, wav, seis_t, seis_rec);
}

void hx_step_f32(const int *g, const float *hs, const float *u, float *U, const float *Cf,
                 const float *ax, const float *bx, const float *az, const float *bz, const float *pXi,
                 float *pXo, float *cX, const float *pZi, float *pZo, float *cZ, const int *recmap,
                 const float *seis_row, float *snap, float *I, float *Is, float *Ic, float *Isc)
{
    run<float>(g, hs, u, U, Cf, ax, bx, az, bz, pXi, pXo, cX, pZi, pZo, cZ, recmap, seis_row, snap, I,
               Is, Ic, Isc);
}

void hx_step_f64(const int *g, const double *hs, const double *u, double *U, const double *Cf,
                 const double *ax, const double *bx, const double *az, const double *bz,
                 const double *pXi, double *pXo, double *cX, const double *pZi, double *pZo,
                 double *cZ, const int *recmap, const float *seis_row, double *snap, double *I,
                 double *Is, double *Ic, double *Isc)
{
    run<double>(g, hs, u, U, Cf, ax, bx, az, bz, pXi, pXo, cX, pZi, pZo, cZ, recmap, seis_row, snap, I,
                Is, Ic, Isc);
}
}
