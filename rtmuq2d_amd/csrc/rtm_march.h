/*
 * rtm_march.h -- the streaming interior kernel body (the part of each grid
 * with no CPML strip), written as phases for the GPU and the CPU harness.
 *
 * A work item {sim, x0, zA, zB} is a 64-column strip of one simulation's
 * interior, rows zA..zB-1 (a whole number of chunks of TZ = 4*RZ rows).  One
 * workgroup marches down it: while chunk c is computed from an LDS ring of u
 * rows, the next chunk's new u rows and its C, U_old, imaging and trace
 * operands are already in flight into registers, so HBM streams
 * continuously; u rows shared by neighbouring chunks (the radius-4 halo) are
 * read from HBM once.
 *
 * Interior points have no CPML term, so a point's update is the stencil
 * (SRC/stencilDFM.c:23-28) followed by source add / receiver replacement and
 * snapshot store / imaging, with the same value forms as rtm_tile.h.
 */
#ifndef RTM_MARCH_H
#define RTM_MARCH_H

#include "rtm_tile.h"

namespace rtm {

template <int RZ>
struct MarchShape {
    static constexpr int TZ = 4 * RZ;
    /* ring rows: the chunk being computed plus its halo (TZ+8) and the next
     * chunk's new rows (TZ), rounded up to a power of two */
    static constexpr int NR = (2 * TZ + 8) <= 32 ? 32 : ((2 * TZ + 8) <= 64 ? 64 : 128);
};

template <typename R, int RZ>
struct MarchLds {
    R u[MarchShape<RZ>::NR][72];   /* row k at [k & (NR-1)], column x0-4+c at [.][c] */
};

/* Operands of one chunk, per thread. */
template <typename R, int RZ>
struct MarchRegs {
    R u0[RZ + 2], u1[RZ + 2];  /* new u rows (RZ of them; RZ+2 in the prologue) */
    R C[RZ], Uold[RZ];
    R sv[RZ], Iv[RZ], Isv[RZ], Icv[RZ], Iscv[RZ];
    int recr;
    R recv;
};

/* Load u rows [k0, k0 + 4*nrows) (nrows per thread) of columns x0-4 .. x0+67. */
template <typename R, int RZ, int NROWS>
RTM_HD void march_fetch_rows(MarchRegs<R, RZ> &G, const SimDesc<R> &S, const StepView<R> &V, int x0,
                             int k0, int tx, int ty)
{
    const int ia = x0 - 4 + tx, ib = x0 + 60 + (tx & 7);
    const int ca = clampi(ia, 0, S.X - 1), cb = clampi(ib, 0, S.X - 1);
#pragma unroll
    for (int t = 0; t < NROWS; t++) {
        const int k = k0 + ty + 4 * t;
        const int base = clampi(k, 0, S.Z - 1) * S.P;
        G.u0[t] = ldg(V.u, base + ca);
        G.u1[t] = ldg(V.u, base + cb);
    }
}

template <typename R, int RZ, int NROWS>
RTM_HD void march_stage_rows(MarchLds<R, RZ> &L, const MarchRegs<R, RZ> &G, int k0, int tx, int ty)
{
    constexpr int NR = MarchShape<RZ>::NR;
#pragma unroll
    for (int t = 0; t < NROWS; t++) {
        const int r = (k0 + ty + 4 * t) & (NR - 1);
        L.u[r][tx] = G.u0[t];
        L.u[r][64 + (tx & 7)] = G.u1[t];
    }
}

/* Operands of the chunk starting at row zc (own points rows zc+ty*RZ+rr). */
template <typename R, int RZ>
RTM_HD void march_fetch_points(MarchRegs<R, RZ> &G, const SimDesc<R> &S, const StepView<R> &V, int x0,
                               int zc, int tx, int ty)
{
    const int i = x0 + tx;
#pragma unroll
    for (int rr = 0; rr < RZ; rr++) {
        const int j = (zc + ty * RZ + rr) * S.P + i;
        G.C[rr] = ldg(S.C, j);
        G.Uold[rr] = ldg(V.U, j);
    }
    if (!V.forward && V.snap) {
        const int ii = clampi(i - S.TL, 0, S.nx - 1);
#pragma unroll
        for (int rr = 0; rr < RZ; rr++) {
            const int kk = clampi(zc + ty * RZ + rr - S.TU, 0, S.nz - 1);
            const int q = kk * S.nx + ii;
            G.sv[rr] = ldg(V.snap, q);
            G.Iv[rr] = ldg(S.I, q);
            G.Isv[rr] = ldg(S.Is, q);
            if (S.flags & F_KAHAN) {
                G.Icv[rr] = ldg(S.Ic, q);
                G.Iscv[rr] = ldg(S.Isc, q);
            }
        }
    }
    G.recr = -1;
    G.recv = R(0);
    const int rrow = S.rec_row - (zc + ty * RZ);
    if (rrow >= 0 && rrow < RZ && (!V.forward || V.rec_row_out)) {   /* uniform */
        G.recr = ldg(S.recmap, i);
        if (!V.forward) {
            const R v = R(ldg(V.seis_row, G.recr > 0 ? G.recr : 0));
            G.recv = G.recr >= 0 ? v : R(0);
        }
    }
}

/* The fused interior update of this thread's RZ points of the chunk at zc. */
template <typename R, int RZ>
RTM_HD void march_points(const MarchLds<R, RZ> &L, const MarchRegs<R, RZ> &G, const SimDesc<R> &S,
                         const StepView<R> &V, int x0, int zc, int tx, int ty)
{
    constexpr int NR = MarchShape<RZ>::NR;
    const int i = x0 + tx;
    const int c = tx + 4;
    const int kb = zc + ty * RZ;
    R col[RZ + 8];
#pragma unroll
    for (int t = 0; t < RZ + 8; t++) col[t] = L.u[(kb - 4 + t) & (NR - 1)][c];
    const bool img_col = i - S.TL >= 0 && i - S.TL < S.nx;
#pragma unroll
    for (int rr = 0; rr < RZ; rr++) {
        const int k = kb + rr;
        const int j = k * S.P + i;
        const R *lu = &L.u[k & (NR - 1)][c];
        const R uc = col[rr + 4];
        R U = stencil_val(G.C[rr], G.Uold[rr], uc, lu[-1], lu[1], col[rr + 3], col[rr + 5], lu[-2], lu[2],
                          col[rr + 2], col[rr + 6], lu[-3], lu[3], col[rr + 1], col[rr + 7], lu[-4], lu[4],
                          col[rr + 0], col[rr + 8]);
        const bool rec_row = k == S.rec_row && G.recr >= 0;
        if (V.forward) {
            if (V.src && j == S.src_j) U = U + V.src_w * G.C[rr];
        } else {
            U = rec_row ? G.recv : U;
        }
        stg(V.U, j, U);
        if (V.forward && rec_row) stg(V.rec_row_out, G.recr, float(U));
        const int kk = k - S.TU;
        if (V.snap && img_col && kk >= 0 && kk < S.nz) {
            const int q = kk * S.nx + (i - S.TL);
            if (V.forward) {
                stg(V.snap, q, U);
            } else if (S.flags & F_KAHAN) {
                R I = G.Iv[rr], Ic = G.Icv[rr], Is = G.Isv[rr], Isc = G.Iscv[rr];
                const R s = G.sv[rr];
                kahan_add(I, Ic, s * U);
                kahan_add(Is, Isc, s * s);
                stg(S.I, q, I); stg(S.Ic, q, Ic); stg(S.Is, q, Is); stg(S.Isc, q, Isc);
            } else {
                const R s = G.sv[rr];
                stg(S.I, q, G.Iv[rr] + s * U);
                stg(S.Is, q, G.Isv[rr] + s * s);
            }
        }
    }
}

template <typename R, int RZ>
RTM_HD void march_copy(MarchRegs<R, RZ> &dst, const MarchRegs<R, RZ> &src)
{
#pragma unroll
    for (int t = 0; t < RZ; t++) {
        dst.u0[t] = src.u0[t]; dst.u1[t] = src.u1[t];
        dst.C[t] = src.C[t]; dst.Uold[t] = src.Uold[t];
        dst.sv[t] = src.sv[t]; dst.Iv[t] = src.Iv[t]; dst.Isv[t] = src.Isv[t];
        dst.Icv[t] = src.Icv[t]; dst.Iscv[t] = src.Iscv[t];
    }
    dst.recr = src.recr;
    dst.recv = src.recv;
}

}  // namespace rtm

#endif /* RTM_MARCH_H */
