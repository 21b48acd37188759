// cpu_step_harness.cpp -- runs the device step arithmetic (rtm_step.h) on the
// CPU so the per-point fused update can be checked against the oracle
// without a GPU.  Built by tests/test_cpu_harness.py with
// g++ -O2 -ffp-contract=off; test infrastructure only.
#include <cstddef>

#include "rtm_step.h"
#include "rtm_tile.h"
#include "rtm_march.h"

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

// Emulation of k_tiles for one simulation at step n: every tile, phase by
// phase, all 256 threads of a phase before the next (the barriers).
template <typename R, int RZ, int CLS>
static void run_tile(rtm::TileLds<R, RZ> &L, rtm::TileRegs<R, RZ> *G, const rtm::SimDesc<R> &S,
                     const rtm::StepView<R> &V, int x0, int z0)
{
    for (int ty = 0; ty < 4; ty++)
        for (int tx = 0; tx < 64; tx++) {
            rtm::tile_fetch<R, RZ, CLS>(G[ty * 64 + tx], S, V, x0, z0, tx, ty);
            rtm::tile_stage<R, RZ, CLS>(L, G[ty * 64 + tx], tx, ty);
        }
    if (CLS)
        for (int ty = 0; ty < 4; ty++)
            for (int tx = 0; tx < 64; tx++) rtm::tile_psi<R, RZ, CLS>(L, G[ty * 64 + tx], S, x0, z0, tx, ty);
    for (int ty = 0; ty < 4; ty++)
        for (int tx = 0; tx < 64; tx++) rtm::tile_points<R, RZ, CLS>(L, G[ty * 64 + tx], S, V, x0, z0, tx, ty);
}

// Emulation of k_march for one work item {x0, zA, zB}.
template <typename R, int RZ>
static void run_march(rtm::MarchLds<R, RZ> &L, rtm::MarchRegs<R, RZ> *cur, rtm::MarchRegs<R, RZ> *nxt,
                      const rtm::SimDesc<R> &S, const rtm::StepView<R> &V, int x0, int zA, int zB)
{
    const int TZ = 4 * RZ;
    for (int th = 0; th < 256; th++) {
        rtm::march_fetch_rows<R, RZ, RZ + 2>(cur[th], S, V, x0, zA - 4, th & 63, th >> 6);
        rtm::march_fetch_points(cur[th], S, V, x0, zA, th & 63, th >> 6);
    }
    for (int th = 0; th < 256; th++) rtm::march_stage_rows<R, RZ, RZ + 2>(L, cur[th], zA - 4, th & 63, th >> 6);
    for (int zc = zA; zc < zB; zc += TZ) {
        const bool more = zc + TZ < zB;
        if (more)
            for (int th = 0; th < 256; th++) {
                rtm::march_fetch_rows<R, RZ, RZ>(nxt[th], S, V, x0, zc + TZ + 4, th & 63, th >> 6);
                rtm::march_fetch_points(nxt[th], S, V, x0, zc + TZ, th & 63, th >> 6);
            }
        for (int th = 0; th < 256; th++) rtm::march_points(L, cur[th], S, V, x0, zc, th & 63, th >> 6);
        if (more)
            for (int th = 0; th < 256; th++) {
                rtm::march_stage_rows<R, RZ, RZ>(L, nxt[th], zc + TZ + 4, th & 63, th >> 6);
                rtm::march_copy(cur[th], nxt[th]);
            }
    }
}

// One step of one simulation: tiles where a CPML strip is touched (all tiles
// when march_rows == 0) and marched interior segments otherwise, split the
// way the library's work-list builder splits them.
template <typename R, int RZ>
static void run_tiles(const rtm::SimDesc<R> &S, int n, int march_rows)
{
    const rtm::StepView<R> V = rtm::step_view(S, n);
    if (!V.active) return;
    auto *L = new rtm::TileLds<R, RZ>();
    auto *G = new rtm::TileRegs<R, RZ>[256];
    auto *M = new rtm::MarchLds<R, RZ>();
    auto *cur = new rtm::MarchRegs<R, RZ>[256];
    auto *nxt = new rtm::MarchRegs<R, RZ>[256];
    const int TZ = 4 * RZ;
    const int seg = march_rows > 0 ? (march_rows + TZ - 1) / TZ * TZ : 0;
    for (int x0 = rtm::kHalf; x0 < S.X - rtm::kHalf; x0 += 64) {
        int zA = -1, zB = -1;
        for (int z0 = rtm::kHalf; z0 < S.Z - rtm::kHalf; z0 += TZ) {
            const int cls = rtm::tile_class(S, x0, z0, TZ);
            if (seg && cls == 0) {
                if (zA < 0) zA = z0;
                zB = z0 + TZ;
                continue;
            }
            // the LDS holds stale data from the previous tile, as on the GPU
            switch (cls) {
            case 0: run_tile<R, RZ, 0>(*L, G, S, V, x0, z0); break;
            case 1: run_tile<R, RZ, 1>(*L, G, S, V, x0, z0); break;
            case 2: run_tile<R, RZ, 2>(*L, G, S, V, x0, z0); break;
            default: run_tile<R, RZ, 3>(*L, G, S, V, x0, z0); break;
            }
        }
        for (int z = zA; zA >= 0 && z < zB; z += seg)
            run_march<R, RZ>(*M, cur, nxt, S, V, x0, z, z + seg < zB ? z + seg : zB);
    }
    delete[] nxt;
    delete[] cur;
    delete M;
    delete[] G;
    delete L;
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
    const int n = g[17], rz = g[18], march_rows = g[19];
    S.h1 = hs[0]; S.h2 = hs[1]; S.h3 = hs[2];
    S.ubuf[0] = ptr[0]; S.ubuf[1] = ptr[1]; S.C = ptr[2];
    S.ax = ptr[3]; S.bx = ptr[4]; S.az = ptr[5]; S.bz = ptr[6];
    S.psiX[0] = ptr[7]; S.psiX[1] = ptr[8]; S.psiZ[0] = ptr[9]; S.psiZ[1] = ptr[10];
    S.csiX = ptr[11]; S.csiZ = ptr[12];
    S.snaps = ptr[13]; S.I = ptr[14]; S.Is = ptr[15]; S.Ic = ptr[16]; S.Isc = ptr[17];
    S.recmap = recmap; S.wav = wav; S.seis_t = seis_t; S.seis_rec = seis_rec;
    switch (rz) {
    case 1: run_tiles<R, 1>(S, n, march_rows); break;
    case 2: run_tiles<R, 2>(S, n, march_rows); break;
    case 8: run_tiles<R, 8>(S, n, march_rows); break;
    default: run_tiles<R, 4>(S, n, march_rows); break;
    }
}

extern "C" {
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
