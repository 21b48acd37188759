/*
 * rtm_step.h -- per-point arithmetic of one RTM time step (device layout).
 *
 * One step of rtmUq2D, in the order the reference performs it:
 *   1. interior leapfrog update           SRC/stencilDFM.c:18-28
 *   2. CPML corrections, L, T, R, B       SRC/absorbingBoundaryCondition.c:171-357
 *   3. source add (forward)               SRC/forward_modeling.c:105-109
 *      or receiver replacement (backward) SRC/backward_modeling.c:106-110
 *   4. snapshot store (forward)           SRC/forward_modeling.c:111-124
 *      or imaging (backward)              SRC/backward_modeling.c:112-120,
 *                                         SRC/image_conditions.c:20-33
 * All of it is fused per point.  Every floating-point expression is written
 * in the association order of the reference and the library is compiled
 * with -ffp-contract=off, so the FP32 instantiation reproduces the
 * reference's FP32 results bit for bit; the FP64 instantiation is the same
 * scheme in double precision (the FP64 oracle run).
 *
 * Device layout: x fastest.  A padded field of X x Z points is stored as
 * f[k*P + i] (i = x index, k = z index, P = row pitch >= X), so a wavefront
 * reads 64 consecutive x points of one depth row.
 *
 * CPML memory variables are full-field arrays in the same layout: psiX/csiX
 * hold the left and right strips (x-derivative variables), psiZ/csiZ the top
 * and bottom strips.  The CPML's second pass needs this step's psi at
 * x+-1..4 (or z+-1..4) (SRC/absorbingBoundaryCondition.c:174-217 separates
 * the passes with a barrier).  Here psi is double-buffered: each point reads
 * psi^n from psi_in, recomputes psi^{n+1} at the 8 neighbours it needs (the
 * identical expression, hence identical bits) and writes only its own
 * psi^{n+1} to psi_out, so no grid-wide barrier is needed inside a step.
 *
 * RTM_HD marks functions that are compiled for the GPU and, in the CPU test
 * harness, for the host.
 */
#ifndef RTM_STEP_H
#define RTM_STEP_H

#include <stddef.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RTM_HD __host__ __device__ __forceinline__
#else
#define RTM_HD static inline
#endif

namespace rtm {

/* 8th-order second-derivative weights, SRC/stencilDFM.h:8-12 (float values). */
constexpr float kA0 = -2.97399944f;
constexpr float kA1 = +1.70507669f;
constexpr float kA2 = -0.25861812f;
constexpr float kA3 = +0.04577745f;
constexpr float kA4 = -0.00523630f;
/* CPML first-difference weights as the reference has them,
 * SRC/absorbingBoundaryCondition.h:21-24 (B3, B4 verbatim). */
constexpr float kB1 = +0.8f;
constexpr float kB2 = -0.2f;
constexpr float kB3 = +0.025396825f;
constexpr float kB4 = -0.001785714f;
constexpr int kHalf = 4;   /* HALFORDERDF */

enum StepMode : int {
    M_STENCIL = 1,      /* interior update                               */
    M_CPML = 2,         /* CPML psi/csi update and corrections           */
    M_SOURCE = 4,       /* fused source add at src_j (n < nw)            */
    M_INJECT = 8,       /* fused receiver replacement on rec_row         */
    M_SNAP = 16,        /* store the interior of U into snap (FP32/FP64) */
    M_IMAGE = 32,       /* I += S*U, Is += S*S over the interior         */
    M_KAHAN = 64,       /* compensated image accumulation                */
};

template <typename R>
struct StepArgs {
    int X, Z, P;              /* padded extent and row pitch           */
    int TL, TR, TU, TD;       /* CPML thicknesses                       */
    int nx, nz;               /* model (image) extent                   */
    int nrec;
    int mode;                 /* StepMode bits                          */
    R h1, h2, h3;             /* 1/h, h*h, 1/(h*h) rounded as the reference */
    const R *u;               /* u^n                                    */
    R *U;                     /* in: u^{n-1}; out: u^{n+1}               */
    const R *C;               /* (v*dt/h)^2                             */
    const R *ax, *bx;         /* full-length x profiles (0 outside L/R) */
    const R *az, *bz;         /* full-length z profiles (0 outside T/B) */
    const R *psiXi; R *psiXo; R *csiX;
    const R *psiZi; R *psiZo; R *csiZ;
    /* source (forward) */
    int src_j;                /* device index of the source point or -1 */
    R src_w;                  /* wavelet sample of this step            */
    /* receivers (backward): U = seis_t[tau*nrec + recmap[i]] on rec_row */
    int rec_row;
    const int *recmap;        /* [X]: receiver index at column i, or -1 */
    const float *seis_row;    /* seis_t + tau*nrec                      */
    /* snapshot store / imaging, interior [nz][nx] x fastest */
    R *snap;                  /* snapshot written (M_SNAP) or read (M_IMAGE) */
    R *I, *Is;                /* image accumulators                      */
    R *Ic, *Isc;              /* Kahan compensations (M_KAHAN)           */
};

/* ---- value forms: every kernel variant evaluates exactly these trees ---- */

/* sum_{m=4..1} B_m (f(+m) - f(-m)), as SRC/absorbingBoundaryCondition.c:183-186;
 * p[m] = f(+m), q[m] = f(-m) */
template <typename R>
RTM_HD R d1_val(R p1, R p2, R p3, R p4, R q1, R q2, R q3, R q4)
{
    return ((R(kB4) * (p4 - q4) + R(kB3) * (p3 - q3)) + R(kB2) * (p2 - q2)) + R(kB1) * (p1 - q1);
}

/* sum_{m=4..1} A_m (f(+m) + f(-m)) + A0 f(0), as :208-212 */
template <typename R>
RTM_HD R d2_val(R c, R p1, R p2, R p3, R p4, R q1, R q2, R q3, R q4)
{
    return (((R(kA4) * (p4 + q4) + R(kA3) * (p3 + q3)) + R(kA2) * (p2 + q2)) + R(kA1) * (p1 + q1))
           + R(kA0) * c;
}

/* Interior update, SRC/stencilDFM.c:23-28; operand order inside each
 * bracket is (x-, x+, z-, z+). */
template <typename R>
RTM_HD R stencil_val(R C, R Uold, R uc, R xm1, R xp1, R zm1, R zp1, R xm2, R xp2, R zm2, R zp2,
                     R xm3, R xp3, R zm3, R zp3, R xm4, R xp4, R zm4, R zp4)
{
    const R s1 = ((xm1 + xp1) + zm1) + zp1;
    const R s2 = ((xm2 + xp2) + zm2) + zp2;
    const R s3 = ((xm3 + xp3) + zm3) + zp3;
    const R s4 = ((xm4 + xp4) + zm4) + zp4;
    const R lap = ((((R(2.0f) * R(kA0)) * uc + R(kA1) * s1) + R(kA2) * s2) + R(kA3) * s3) + R(kA4) * s4;
    return (C * lap + R(2.0f) * uc) - Uold;
}

/* psi^{n+1} = b psi^n + (a D1(u)) / h, as :180-187 */
template <typename R>
RTM_HD R psi_val(R a, R b, R psi_old, R d1, R h1)
{
    return b * psi_old + (a * d1) * h1;
}

/* csi^{n+1} = b csi^n + a (D2(u) / h^2 + dpsi), as :205-213 */
template <typename R>
RTM_HD R csi_val(R a, R b, R csi_old, R d2, R dpsi, R h3)
{
    return b * csi_old + a * (d2 * h3 + dpsi);
}

/* (C h^2)(dpsi + csi), the correction added to U, as :215 */
template <typename R>
RTM_HD R corr_val(R C, R h2, R dpsi, R cs)
{
    return (C * h2) * (dpsi + cs);
}

/* ---- pointer forms over a strided array ---- */

template <typename R>
RTM_HD R first_diff(const R *f, ptrdiff_t j, ptrdiff_t s)
{
    return d1_val(f[j + s], f[j + 2 * s], f[j + 3 * s], f[j + 4 * s],
                  f[j - s], f[j - 2 * s], f[j - 3 * s], f[j - 4 * s]);
}

template <typename R>
RTM_HD R second_diff(const R *f, ptrdiff_t j, ptrdiff_t s)
{
    return d2_val(f[j], f[j + s], f[j + 2 * s], f[j + 3 * s], f[j + 4 * s],
                  f[j - s], f[j - 2 * s], f[j - 3 * s], f[j - 4 * s]);
}

template <typename R>
RTM_HD R stencil_point(const R *u, R C, R Uold, ptrdiff_t j, ptrdiff_t P)
{
    return stencil_val(C, Uold, u[j], u[j - 1], u[j + 1], u[j - P], u[j + P],
                       u[j - 2], u[j + 2], u[j - 2 * P], u[j + 2 * P],
                       u[j - 3], u[j + 3], u[j - 3 * P], u[j + 3 * P],
                       u[j - 4], u[j + 4], u[j - 4 * P], u[j + 4 * P]);
}

template <typename R>
RTM_HD R psi_next(R a, R b, R psi_old, const R *u, ptrdiff_t j, ptrdiff_t s, R h1)
{
    return psi_val(a, b, psi_old, first_diff(u, j, s), h1);
}

template <typename R>
RTM_HD bool in_x_strip(const StepArgs<R> &A, int i)
{
    return (i >= kHalf && i < A.TL - kHalf) || (i >= A.X - A.TR + kHalf && i < A.X - kHalf);
}

template <typename R>
RTM_HD bool in_z_strip(const StepArgs<R> &A, int k)
{
    return (k >= kHalf && k < A.TU - kHalf) || (k >= A.Z - A.TD + kHalf && k < A.Z - kHalf);
}

/* CPML correction of one axis at a strip point (i, k), device index j.
 * `axis` 0 = x (left/right strips), 1 = z (top/bottom strips).  Writes this
 * point's psi^{n+1} and csi^{n+1}; returns (C h^2)(dpsi + csi^{n+1}),
 * SRC/absorbingBoundaryCondition.c:198-215. */
template <typename R>
RTM_HD R cpml_correction(const StepArgs<R> &A, int i, int k, ptrdiff_t j, int axis)
{
    const ptrdiff_t s = axis == 0 ? 1 : (ptrdiff_t)A.P;
    const int c = axis == 0 ? i : k;                 /* coordinate along the strip normal */
    const R *a = axis == 0 ? A.ax : A.az;
    const R *b = axis == 0 ? A.bx : A.bz;
    const R *psi_in = axis == 0 ? A.psiXi : A.psiZi;
    R *psi_out = axis == 0 ? A.psiXo : A.psiZo;
    R *csi = axis == 0 ? A.csiX : A.csiZ;

    // psi^{n+1} at the 8 neighbours along the normal: recomputed where the
    // reference's first pass updates psi, the stored value elsewhere (it is
    // never written there; zero in a modeling run)
    R pm[5], pp[5];
#pragma unroll
    for (int m = 1; m <= 4; m++) {
        const int cm = c - m, cp = c + m;
        const bool okm = axis == 0 ? in_x_strip(A, cm) : in_z_strip(A, cm);
        const bool okp = axis == 0 ? in_x_strip(A, cp) : in_z_strip(A, cp);
        pm[m] = okm ? psi_next(a[cm], b[cm], psi_in[j - m * s], A.u, j - m * s, s, A.h1) : psi_in[j - m * s];
        pp[m] = okp ? psi_next(a[cp], b[cp], psi_in[j + m * s], A.u, j + m * s, s, A.h1) : psi_in[j + m * s];
    }
    psi_out[j] = psi_next(a[c], b[c], psi_in[j], A.u, j, s, A.h1);
    const R dpsi = d1_val(pp[1], pp[2], pp[3], pp[4], pm[1], pm[2], pm[3], pm[4]) * A.h1;
    const R cs = csi_val(a[c], b[c], csi[j], second_diff(A.u, j, s), dpsi, A.h3);
    csi[j] = cs;
    return corr_val(A.C[j], A.h2, dpsi, cs);
}

/* Top-strip correction in the outer columns i < 4 and i >= X-4, 4 <= k < TU-4.
 * The reference's second top pass runs over the full width
 * (SRC/absorbingBoundaryCondition.c:239) while its first pass updates psi only
 * for 4 <= i < X-4 (:221), so here dpsi uses the stored psi.  In a modeling
 * run these columns hold zeros and the correction is exactly zero; it matters
 * only for arbitrary state (the per-step entry point) or for injection points
 * placed in those columns. */
template <typename R>
RTM_HD void cpml_top_edge_point(const StepArgs<R> &A, int i, int k)
{
    const ptrdiff_t s = (ptrdiff_t)A.P, j = (ptrdiff_t)k * A.P + i;
    const R dpsi = first_diff(A.psiZi, j, s) * A.h1;
    const R cs = csi_val(A.az[k], A.bz[k], A.csiZ[j], second_diff(A.u, j, s), dpsi, A.h3);
    A.csiZ[j] = cs;
    A.U[j] = A.U[j] + corr_val(A.C[j], A.h2, dpsi, cs);
}

/* Kahan-compensated acc += v (fp-contract is off, so the compensation
 * survives compilation). */
template <typename R>
RTM_HD void kahan_add(R &acc, R &comp, R v)
{
    const R y = v - comp;
    const R t = acc + y;
    comp = (t - acc) - y;
    acc = t;
}

/* The fused update of one point (i, k) with 4 <= i < X-4, 4 <= k < Z-4. */
template <typename R>
RTM_HD void step_point(const StepArgs<R> &A, int i, int k)
{
    const ptrdiff_t j = (ptrdiff_t)k * A.P + i;
    R U = A.U[j];
    if (A.mode & M_STENCIL) U = stencil_point(A.u, A.C[j], U, j, (ptrdiff_t)A.P);
    if (A.mode & M_CPML) {
        const bool inL = i < A.TL - kHalf, inR = i >= A.X - A.TR + kHalf;
        const bool inT = k < A.TU - kHalf, inB = k >= A.Z - A.TD + kHalf;
        if (inL) U = U + cpml_correction(A, i, k, j, 0);
        if (inT) U = U + cpml_correction(A, i, k, j, 1);
        if (inR) U = U + cpml_correction(A, i, k, j, 0);
        if (inB) U = U + cpml_correction(A, i, k, j, 1);
    }
    if ((A.mode & M_SOURCE) && j == A.src_j) U = U + A.src_w * A.C[j];
    if ((A.mode & M_INJECT) && k == A.rec_row) {
        const int r = A.recmap[i];
        if (r >= 0) U = R(A.seis_row[r]);
    }
    A.U[j] = U;
    if (A.mode & (M_SNAP | M_IMAGE)) {
        const int ii = i - A.TL, kk = k - A.TU;
        if (ii >= 0 && ii < A.nx && kk >= 0 && kk < A.nz) {
            const ptrdiff_t q = (ptrdiff_t)kk * A.nx + ii;
            if (A.mode & M_SNAP) A.snap[q] = U;
            if (A.mode & M_IMAGE) {
                const R s = A.snap[q];
                if (A.mode & M_KAHAN) {
                    R I = A.I[q], Ic = A.Ic[q], Is = A.Is[q], Isc = A.Isc[q];
                    kahan_add(I, Ic, s * U);
                    kahan_add(Is, Isc, s * s);
                    A.I[q] = I; A.Ic[q] = Ic; A.Is[q] = Is; A.Isc[q] = Isc;
                } else {
                    A.I[q] = A.I[q] + s * U;
                    A.Is[q] = A.Is[q] + s * s;
                }
            }
        }
    }
}

}  // namespace rtm

#endif /* RTM_STEP_H */
