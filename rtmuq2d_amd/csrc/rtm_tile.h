/*
 * rtm_tile.h -- the LDS-tiled step kernel body, written as barrier-separated
 * phases so the same code runs on the GPU (one workgroup per tile) and in the
 * CPU test harness (phases run one after another for all 256 threads).
 *
 * A launch advances any number of simulations ("sims") by one time step.
 * Each sim is described by a SimDesc in device memory (its grid may differ
 * from the others'), and the launch walks a flat list of tiles
 * {sim, x0, z0, class}.  A tile covers 64 x TZ points (TZ = 4*RZ) of the
 * stepped region 4 <= i < X-4, 4 <= k < Z-4; 256 threads = 64 columns x 4
 * row groups, each thread owning RZ consecutive rows of one column.
 *
 * Phase 1  load u^n over the tile plus a halo into LDS: radius 4, or radius
 *          8 across a CPML strip (the strip's psi is needed 4 points out,
 *          and psi itself reads u 4 points out);
 * Phase 2  (strip tiles) psi^{n+1} over the tile plus a radius-4 halo into
 *          LDS, once per point -- replacing the reference's separate first
 *          pass and barrier (SRC/absorbingBoundaryCondition.c:174-217);
 * Phase 3  per point: stencil, CPML corrections in the order L, T, R, B,
 *          source add or receiver replacement, snapshot store or imaging.
 *
 * The arithmetic is the value forms of rtm_step.h, so results are
 * bit-identical to rtm::step_point (and to the reference's FP32 arithmetic).
 * psi^{n+1} of a point outside the update range is 0: in a modeling pass
 * those entries are never written and start at zero.
 */
#ifndef RTM_TILE_H
#define RTM_TILE_H

#include "rtm_step.h"

namespace rtm {

enum SimPhase : int { PH_IDLE = 0, PH_FORWARD = 1, PH_BACKWARD = 2 };
enum SimFlags : int {
    F_KAHAN = 1,    /* compensated image accumulation          */
    F_RECORD = 2,   /* fused seismogram recording (forward)     */
    F_SNAP = 4,     /* store snapshots (forward)                */
    F_IMAGE = 8,    /* image against snapshots (backward)       */
};
enum TileClass : int { T_XSTRIP = 1, T_ZSTRIP = 2 };

template <typename R>
struct SimDesc {
    int X, Z, P;
    int TL, TR, TU, TD;
    int nx, nz;
    int nt, TS, nw, nrec;
    int phase;              /* SimPhase */
    int flags;              /* SimFlags */
    int src_j;              /* device index of the source, -1 if none */
    int rec_row;
    R h1, h2, h3;
    R *ubuf[2];             /* u at step n is ubuf[n & 1]              */
    const R *C;
    const R *ax, *bx, *az, *bz;
    R *psiX[2], *psiZ[2];   /* psi^n is psi?[n & 1]                    */
    R *csiX, *csiZ;
    const int *recmap;      /* [X] receiver at column i on rec_row, -1 */
    const float *wav;       /* [nw]                                    */
    const float *seis_t;    /* [nt][nrec] input traces (backward)      */
    float *seis_rec;        /* [nt][nrec] recorded traces (forward)    */
    R *snaps;               /* [nt/TS][nz][nx]                          */
    R *I, *Is, *Ic, *Isc;   /* [nz][nx]                                 */
};

template <typename R, int RZ>
struct TileLds {
    static constexpr int TZ = 4 * RZ;
    R u[TZ + 16][80];       /* rows z0-8 .. z0+TZ+7, columns x0-8 .. x0+71 */
    R px[TZ][72];           /* psiX^{n+1}: rows z0 .., columns x0-4 .. x0+67 */
    R pz[TZ + 8][64];       /* psiZ^{n+1}: rows z0-4 .., columns x0 .. x0+63 */
};

/* Per-launch view of one sim at step n. */
template <typename R>
struct StepView {
    int n, active, forward;
    const R *u;
    R *U;
    const R *psiXi, *psiZi;
    R *psiXo, *psiZo;
    int src;                /* add the source this step */
    R src_w;
    const float *seis_row;  /* backward: traces at tau */
    float *rec_row_out;     /* forward recording row   */
    R *snap;                /* snapshot written (forward) or read (backward) */
};

template <typename R>
RTM_HD StepView<R> step_view(const SimDesc<R> &S, int n)
{
    StepView<R> v{};
    v.n = n;
    v.active = S.phase != PH_IDLE && n >= 0 && n < S.nt;
    if (!v.active) return v;
    v.forward = S.phase == PH_FORWARD;
    v.u = S.ubuf[n & 1];
    v.U = S.ubuf[(n + 1) & 1];
    v.psiXi = S.psiX[n & 1];
    v.psiXo = S.psiX[(n + 1) & 1];
    v.psiZi = S.psiZ[n & 1];
    v.psiZo = S.psiZ[(n + 1) & 1];
    const size_t nint = (size_t)S.nx * S.nz;
    if (v.forward) {
        v.src = n < S.nw && S.src_j >= 0;
        v.src_w = v.src ? R(S.wav[n]) : R(0);
        v.rec_row_out = (S.flags & F_RECORD) ? S.seis_rec + (size_t)n * S.nrec : nullptr;
        v.snap = ((S.flags & F_SNAP) && n % S.TS == 0) ? S.snaps + (size_t)(n / S.TS) * nint : nullptr;
    } else {
        const int tau = S.nt - 1 - n;
        v.seis_row = S.seis_t + (size_t)tau * S.nrec;
        v.snap = ((S.flags & F_IMAGE) && tau % S.TS == 0) ? S.snaps + (size_t)(tau / S.TS) * nint : nullptr;
    }
    return v;
}

template <typename R>
RTM_HD bool x_strip(const SimDesc<R> &S, int i)
{
    return (i >= kHalf && i < S.TL - kHalf) || (i >= S.X - S.TR + kHalf && i < S.X - kHalf);
}

template <typename R>
RTM_HD bool z_strip(const SimDesc<R> &S, int k)
{
    return (k >= kHalf && k < S.TU - kHalf) || (k >= S.Z - S.TD + kHalf && k < S.Z - kHalf);
}

/* Class of the tile with origin (x0, z0): which strips its points touch. */
template <typename R>
RTM_HD int tile_class(const SimDesc<R> &S, int x0, int z0, int TZ)
{
    const int x1 = x0 + 64 < S.X - kHalf ? x0 + 64 : S.X - kHalf;   /* exclusive */
    const int z1 = z0 + TZ < S.Z - kHalf ? z0 + TZ : S.Z - kHalf;
    int cls = 0;
    if (x0 < S.TL - kHalf || x1 > S.X - S.TR + kHalf) cls |= T_XSTRIP;
    if (z0 < S.TU - kHalf || z1 > S.Z - S.TD + kHalf) cls |= T_ZSTRIP;
    return cls;
}

/* Per-thread values fetched from HBM before the first barrier, so that the
 * whole step costs one memory round trip: the thread's share of the u tile,
 * and the C, U_old, csi, psi and profile values its later phases need. */
template <typename R, int RZ>
struct TileRegs {
    R u0[RZ + 4], u1[RZ + 4];   /* LDS rows ty+4t, columns tx and 64+tx */
    R C[RZ], Uold[RZ];          /* own points: rows ty*RZ+rr, column tx */
    R csx[RZ], csz[RZ];
    R px0[RZ], px1[RZ];         /* psiX^n at the psi-phase points       */
    R pz[RZ + 2];               /* psiZ^n at the psi-phase points       */
    R ax0, bx0, ax1, bx1;       /* x profiles at psi columns tx-4, tx+60 */
    R axo, bxo;                 /* x profiles at the own column         */
    R sv[RZ], Iv[RZ], Isv[RZ];  /* imaging: snapshot and image accumulators */
    R Icv[RZ], Iscv[RZ];        /* Kahan compensations                  */
    int recr;                   /* receiver at the own column (rec_row), or -1 */
    R recv;                     /* its trace sample (backward)          */
};

/* Phase 1a: issue every global load of the step.  tx in [0,64), ty in [0,4). */
template <typename R, int RZ>
RTM_HD void tile_fetch(TileRegs<R, RZ> &G, const SimDesc<R> &S, const StepView<R> &V, int x0, int z0,
                       int cls, int tx, int ty)
{
    constexpr int TZ = 4 * RZ;
    const int hx = (cls & T_XSTRIP) ? 8 : 4, hz = (cls & T_ZSTRIP) ? 8 : 4;
    const int i0 = x0 - 8 + tx, i1 = x0 + 56 + tx;          /* LDS columns tx and 64+tx */
    const bool c0 = tx >= 8 - hx && i0 >= 0 && i0 < S.X;
    const bool c1 = tx < 8 + hx && i1 < S.X;
#pragma unroll
    for (int t = 0; t < RZ + 4; t++) {
        const int r = ty + 4 * t, k = z0 - 8 + r;
        const bool rok = r >= 8 - hz && r < TZ + 8 + hz && k >= 0 && k < S.Z;
        const ptrdiff_t base = (ptrdiff_t)k * S.P;
        G.u0[t] = (rok && c0) ? V.u[base + i0] : R(0);
        G.u1[t] = (rok && c1) ? V.u[base + i1] : R(0);
    }
    const int i = x0 + tx;
    const bool col_ok = i < S.X - kHalf;
    const bool inX = col_ok && (i < S.TL - kHalf || i >= S.X - S.TR + kHalf);
    G.axo = inX ? S.ax[i] : R(0);
    G.bxo = inX ? S.bx[i] : R(0);
#pragma unroll
    for (int rr = 0; rr < RZ; rr++) {
        const int k = z0 + ty * RZ + rr;
        const bool ok = col_ok && k < S.Z - kHalf;
        const ptrdiff_t j = (ptrdiff_t)k * S.P + i;
        G.C[rr] = ok ? S.C[j] : R(0);
        G.Uold[rr] = ok ? V.U[j] : R(0);
        G.csx[rr] = (ok && inX) ? S.csiX[j] : R(0);
        const bool inZ = ok && (k < S.TU - kHalf || k >= S.Z - S.TD + kHalf);
        G.csz[rr] = inZ ? S.csiZ[j] : R(0);
        /* imaging operands (backward steps that image) */
        const int ii = i - S.TL, kk = k - S.TU;
        const bool img = ok && V.snap && !V.forward && ii >= 0 && ii < S.nx && kk >= 0 && kk < S.nz;
        const ptrdiff_t q = (ptrdiff_t)kk * S.nx + ii;
        G.sv[rr] = img ? V.snap[q] : R(0);
        G.Iv[rr] = img ? S.I[q] : R(0);
        G.Isv[rr] = img ? S.Is[q] : R(0);
        const bool kah = img && (S.flags & F_KAHAN);
        G.Icv[rr] = kah ? S.Ic[q] : R(0);
        G.Iscv[rr] = kah ? S.Isc[q] : R(0);
    }
    /* receiver on the own column, if the receiver row is one of the own rows */
    const int rrow = S.rec_row - (z0 + ty * RZ);
    const bool has_rec = col_ok && rrow >= 0 && rrow < RZ && S.rec_row < S.Z - kHalf &&
                         (!V.forward || V.rec_row_out);
    G.recr = has_rec ? S.recmap[i] : -1;
    G.recv = (!V.forward && G.recr >= 0) ? R(V.seis_row[G.recr]) : R(0);
    if (cls & T_XSTRIP) {
        const int ia = x0 - 4 + tx, ib = x0 + 60 + tx;     /* psi columns c = tx, 64+tx */
        const bool oka = x_strip(S, ia), okb = tx < 8 && x_strip(S, ib);
        G.ax0 = oka ? S.ax[ia] : R(0);
        G.bx0 = oka ? S.bx[ia] : R(0);
        G.ax1 = okb ? S.ax[ib] : R(0);
        G.bx1 = okb ? S.bx[ib] : R(0);
#pragma unroll
        for (int t = 0; t < RZ; t++) {
            const int k = z0 + ty + 4 * t;
            const bool kok = k < S.Z - kHalf;
            G.px0[t] = (kok && oka) ? V.psiXi[(ptrdiff_t)k * S.P + ia] : R(0);
            G.px1[t] = (kok && okb) ? V.psiXi[(ptrdiff_t)k * S.P + ib] : R(0);
        }
    }
    if (cls & T_ZSTRIP) {
#pragma unroll
        for (int t = 0; t < RZ + 2; t++) {
            const int k = z0 - 4 + ty + 4 * t;
            G.pz[t] = (col_ok && z_strip(S, k)) ? V.psiZi[(ptrdiff_t)k * S.P + i] : R(0);
        }
    }
}

/* Phase 1b: the fetched u into LDS (rows z0-8 .., columns x0-8 ..). */
template <typename R, int RZ>
RTM_HD void tile_stage(TileLds<R, RZ> &L, const TileRegs<R, RZ> &G, int tx, int ty)
{
#pragma unroll
    for (int t = 0; t < RZ + 4; t++) {
        L.u[ty + 4 * t][tx] = G.u0[t];
        if (tx < 16) L.u[ty + 4 * t][64 + tx] = G.u1[t];
    }
}

/* Phase 2 (strip tiles): psi^{n+1} over the tile plus a radius-4 halo. */
template <typename R, int RZ>
RTM_HD void tile_psi(TileLds<R, RZ> &L, const TileRegs<R, RZ> &G, const SimDesc<R> &S, int x0, int z0,
                     int cls, int tx, int ty)
{
    if (cls & T_XSTRIP) {
        const bool oka = x_strip(S, x0 - 4 + tx), okb = tx < 8 && x_strip(S, x0 + 60 + tx);
#pragma unroll
        for (int t = 0; t < RZ; t++) {
            const int r = ty + 4 * t;
            const bool kok = z0 + r < S.Z - kHalf;
            {
                const R *lu = &L.u[r + 8][tx + 4];
                R v = R(0);
                if (kok && oka) {
                    const R d1 = d1_val(lu[1], lu[2], lu[3], lu[4], lu[-1], lu[-2], lu[-3], lu[-4]);
                    v = psi_val(G.ax0, G.bx0, G.px0[t], d1, S.h1);
                }
                L.px[r][tx] = v;
            }
            if (tx < 8) {
                const R *lu = &L.u[r + 8][tx + 68];
                R v = R(0);
                if (kok && okb) {
                    const R d1 = d1_val(lu[1], lu[2], lu[3], lu[4], lu[-1], lu[-2], lu[-3], lu[-4]);
                    v = psi_val(G.ax1, G.bx1, G.px1[t], d1, S.h1);
                }
                L.px[r][64 + tx] = v;
            }
        }
    }
    if (cls & T_ZSTRIP) {
        const int i = x0 + tx;
        const int lc = tx + 8;
#pragma unroll
        for (int t = 0; t < RZ + 2; t++) {
            const int r = ty + 4 * t;
            const int k = z0 - 4 + r;
            R v = R(0);
            if (i < S.X - kHalf && z_strip(S, k)) {
                const int lr = r + 4;
                const R d1 = d1_val(L.u[lr + 1][lc], L.u[lr + 2][lc], L.u[lr + 3][lc], L.u[lr + 4][lc],
                                    L.u[lr - 1][lc], L.u[lr - 2][lc], L.u[lr - 3][lc], L.u[lr - 4][lc]);
                v = psi_val(S.az[k], S.bz[k], G.pz[t], d1, S.h1);
            }
            L.pz[r][tx] = v;
        }
    }
}

/* Phase 3: the fused update of this thread's RZ points. */
template <typename R, int RZ>
RTM_HD void tile_points(const TileLds<R, RZ> &L, const TileRegs<R, RZ> &G, const SimDesc<R> &S,
                        const StepView<R> &V, int x0, int z0, int cls, int tx, int ty)
{
    const int i = x0 + tx;
    if (i >= S.X - kHalf) return;
    const int lc = tx + 8;
    /* this column's u from row ty*RZ-4 to ty*RZ+RZ+3 (tile-relative) */
    R col[RZ + 8];
#pragma unroll
    for (int t = 0; t < RZ + 8; t++) col[t] = L.u[ty * RZ + 4 + t][lc];
    const bool inL = i < S.TL - kHalf, inR = i >= S.X - S.TR + kHalf;
#pragma unroll
    for (int rr = 0; rr < RZ; rr++) {
        const int kl = ty * RZ + rr;         /* tile-relative row */
        const int k = z0 + kl;
        if (k >= S.Z - kHalf) break;
        const ptrdiff_t j = (ptrdiff_t)k * S.P + i;
        const R *lu = &L.u[kl + 8][lc];
        const R uc = col[rr + 4];
        const R xm1 = lu[-1], xp1 = lu[1], xm2 = lu[-2], xp2 = lu[2];
        const R xm3 = lu[-3], xp3 = lu[3], xm4 = lu[-4], xp4 = lu[4];
        const R zm1 = col[rr + 3], zp1 = col[rr + 5], zm2 = col[rr + 2], zp2 = col[rr + 6];
        const R zm3 = col[rr + 1], zp3 = col[rr + 7], zm4 = col[rr + 0], zp4 = col[rr + 8];
        const R Cj = G.C[rr];
        R U = stencil_val(Cj, G.Uold[rr], uc, xm1, xp1, zm1, zp1, xm2, xp2, zm2, zp2, xm3, xp3, zm3, zp3,
                          xm4, xp4, zm4, zp4);
        if (cls) {
            const bool inT = k < S.TU - kHalf, inB = k >= S.Z - S.TD + kHalf;
            R corrX = R(0), corrZ = R(0);
            if (inL || inR) {
                const R *p = &L.px[kl][tx + 4];
                V.psiXo[j] = p[0];
                const R dpsi = d1_val(p[1], p[2], p[3], p[4], p[-1], p[-2], p[-3], p[-4]) * S.h1;
                const R d2 = d2_val(uc, xp1, xp2, xp3, xp4, xm1, xm2, xm3, xm4);
                const R cs = csi_val(G.axo, G.bxo, G.csx[rr], d2, dpsi, S.h3);
                S.csiX[j] = cs;
                corrX = corr_val(Cj, S.h2, dpsi, cs);
            }
            if (inT || inB) {
                const int pr = kl + 4;
                V.psiZo[j] = L.pz[pr][tx];
                const R dpsi = d1_val(L.pz[pr + 1][tx], L.pz[pr + 2][tx], L.pz[pr + 3][tx], L.pz[pr + 4][tx],
                                      L.pz[pr - 1][tx], L.pz[pr - 2][tx], L.pz[pr - 3][tx], L.pz[pr - 4][tx])
                               * S.h1;
                const R d2 = d2_val(uc, zp1, zp2, zp3, zp4, zm1, zm2, zm3, zm4);
                const R cs = csi_val(S.az[k], S.bz[k], G.csz[rr], d2, dpsi, S.h3);
                S.csiZ[j] = cs;
                corrZ = corr_val(Cj, S.h2, dpsi, cs);
            }
            /* additions in the reference's strip order: L, T, R, B */
            if (inL) U = U + corrX;
            if (inT) U = U + corrZ;
            if (inR) U = U + corrX;
            if (inB) U = U + corrZ;
        }
        const bool rec_here = k == S.rec_row && G.recr >= 0;
        if (V.forward) {
            if (V.src && j == S.src_j) U = U + V.src_w * Cj;
        } else if (rec_here) {
            U = G.recv;
        }
        V.U[j] = U;
        if (V.forward && rec_here) V.rec_row_out[G.recr] = float(U);
        if (V.snap) {
            const int ii = i - S.TL, kk = k - S.TU;
            if (ii >= 0 && ii < S.nx && kk >= 0 && kk < S.nz) {
                const ptrdiff_t q = (ptrdiff_t)kk * S.nx + ii;
                if (V.forward) {
                    V.snap[q] = U;
                } else {
                    const R s = G.sv[rr];
                    if (S.flags & F_KAHAN) {
                        R I = G.Iv[rr], Ic = G.Icv[rr], Is = G.Isv[rr], Isc = G.Iscv[rr];
                        kahan_add(I, Ic, s * U);
                        kahan_add(Is, Isc, s * s);
                        S.I[q] = I; S.Ic[q] = Ic; S.Is[q] = Is; S.Isc[q] = Isc;
                    } else {
                        S.I[q] = G.Iv[rr] + s * U;
                        S.Is[q] = G.Isv[rr] + s * s;
                    }
                }
            }
        }
    }
}

}  // namespace rtm

#endif /* RTM_TILE_H */
