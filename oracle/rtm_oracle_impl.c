/*
 * rtm_oracle_impl.c -- CPU restatement of rtmUq2D's RTM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * This file is compiled twice (see oracle/Makefile): with REAL=float and
 * SFX=_f32 (the reference's own FP32 arithmetic) and with REAL=double and
 * SFX=_f64 (the same scheme in FP64).  Both builds use -ffp-contract=off, so
 * each expression below is evaluated in exactly the order written, and that
 * order is the order in which the reference writes it.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - bit-exact against the reference's own C compiled in place from
 *     /root/reference/SRC with -ffp-contract=off (oracle/build_ref.sh), per
 *     function and end to end (rtm_routine);
 *   - closed forms: constant fields, the 1.99764 quadratic factor, CPML
 *     profile formulas evaluated independently in numpy, delta tests for
 *     source/receiver indexing, I = sum S*R for constant fields; the Kahan
 *     image option (no counterpart in the reference) against exactly
 *     rounded sums of constant products, where plain FP32 summation drifts.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "rtm_oracle.h"

#ifndef REAL
#define REAL float
#define SFX _f32
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* 8th-order second-derivative weights, SRC/stencilDFM.h:8-12.  They are float
 * literals in the reference; the FP64 build promotes the same float values. */
#define A0 ((REAL)(-2.97399944f))
#define A1 ((REAL)(+1.70507669f))
#define A2 ((REAL)(-0.25861812f))
#define A3 ((REAL)(+0.04577745f))
#define A4 ((REAL)(-0.00523630f))
/* First-derivative weights as the reference has them,
 * SRC/absorbingBoundaryCondition.h:21-24 (B3, B4 kept verbatim). */
#define B1 ((REAL)(+0.8f))
#define B2 ((REAL)(-0.2f))
#define B3 ((REAL)(+0.025396825f))
#define B4 ((REAL)(-0.001785714f))

#define HALF 4      /* HALFORDERDF, SRC/stencilDFM.h:5 */
#define ORDER 8     /* ORDERDF, SRC/absorbingBoundaryCondition.h:8 */

static inline int dimX(const orc_params *p) { return p->t_left + p->nx + p->t_right; }
static inline int dimZ(const orc_params *p) { return p->t_up + p->nz + p->t_down; }

/* coef = (v*dt/h)^2 on the padded velocity: SRC/forward_modeling.c:71-81,
 * SRC/backward_modeling.c:232-238. */
void FN(orc_coef)(const orc_params *p, REAL *coef, const float *vpad)
{
    const int n = dimX(p) * dimZ(p);
    const REAL aux = (REAL)p->dt / (REAL)p->h;
    for (int j = 0; j < n; j++) {
        const REAL v = (REAL)vpad[j];
        coef[j] = (v * aux) * (v * aux);
    }
}

/* CPML damping profiles: SRC/absorbingBoundaryCondition.c:11-148.
 * Note the hard-coded vmax=4000 and fcut=60 (lines 42-43) and that the left
 * thickness sets sigma_max for every side (line 65). */
void FN(orc_profiles)(const orc_params *p, REAL *aL, REAL *aR, REAL *aT, REAL *aB,
                      REAL *bL, REAL *bR, REAL *bT, REAL *bB)
{
#if defined(ORACLE_DOUBLE)
    const double pi = M_PI;
    const double att = 3.0 * pi;
    const double vmax = 4000.0, fcut = 60.0, npow = 2.0;
    const double h = p->h, dt = p->dt;
    const double fdom = fcut / (3.0 * sqrt(pi));
    const double LL = (p->t_left - ORDER) * h, LR = (p->t_right - ORDER) * h;
    const double LT = (p->t_up - ORDER) * h, LB = (p->t_down - ORDER) * h;
    const double smax = att * (((npow + 1.0) * vmax) / LL);
    const double amax = pi * fdom;
#define POWE(x) exp(x)
#define POW2(x) pow((x), npow)
#else
    const float pi = M_PI;
    const float att = 3.0f * pi;
    const float vmax = 4000.0f, fcut = 60.0f, npow = 2.0f;
    const float h = p->h, dt = p->dt;
    const float fdom = fcut / (3.0f * sqrt(pi));
    const float LL = (p->t_left - ORDER) * h, LR = (p->t_right - ORDER) * h;
    const float LT = (p->t_up - ORDER) * h, LB = (p->t_down - ORDER) * h;
    const float smax = att * (((npow + 1.0f) * vmax) / LL);
    const float amax = pi * fdom;
#define POWE(x) powf((float)M_E, (x))
#define POW2(x) powf((x), npow)
#endif
    const int X = dimX(p), Z = dimZ(p);
    memset(aL, 0, sizeof(REAL) * p->t_left);  memset(bL, 0, sizeof(REAL) * p->t_left);
    memset(aR, 0, sizeof(REAL) * p->t_right); memset(bR, 0, sizeof(REAL) * p->t_right);
    memset(aT, 0, sizeof(REAL) * p->t_up);    memset(bT, 0, sizeof(REAL) * p->t_up);
    memset(aB, 0, sizeof(REAL) * p->t_down);  memset(bB, 0, sizeof(REAL) * p->t_down);

    for (int i = HALF; i < p->t_left - HALF; i++) {          /* :75-86 */
        REAL d = (LL - h * (i - HALF)) / LL;
        REAL s = smax * POW2(d);
        REAL a = amax * (1 - d) + (REAL)0.1f * amax;
        bL[i] = POWE(-(s + a) * dt);
        aL[i] = s * (bL[i] - 1) / (s + a);
    }
    for (int ii = 0, i = X - (p->t_right - HALF); i < X - HALF; i++, ii++) {  /* :90-103 */
        REAL d = (h + h * ii) / LR;
        REAL s = smax * POW2(d);
        REAL a = amax * (1 - d) + (REAL)0.1f * amax;
        bR[ii + HALF] = POWE(-(s + a) * dt);
        aR[ii + HALF] = s * (bR[ii + HALF] - 1) / (s + a);
    }
    for (int k = HALF; k < p->t_up - HALF; k++) {            /* :107-118 */
        REAL d = (LT - h * (k - HALF)) / LT;
        REAL s = smax * POW2(d);
        REAL a = amax * (1 - d) + (REAL)0.1f * amax;
        bT[k] = POWE(-(s + a) * dt);
        aT[k] = s * (bT[k] - 1) / (s + a);
    }
    for (int kk = 0, k = Z - (p->t_down - HALF); k < Z - HALF; k++, kk++) {   /* :122-135 */
        REAL d = (h + h * kk) / LB;
        REAL s = smax * POW2(d);
        REAL a = amax * (1 - d) + (REAL)0.1f * amax;
        bB[kk + HALF] = POWE(-(s + a) * dt);
        aB[kk + HALF] = s * (bB[kk + HALF] - 1) / (s + a);
    }
#undef POWE
#undef POW2
}

/* Interior leapfrog update, in place on U: SRC/stencilDFM.c:7-31. */
void FN(orc_stencil)(const orc_params *p, const REAL *C, const REAL *u, REAL *U)
{
    const int X = dimX(p), Z = dimZ(p);
    const REAL c0 = (REAL)2.0f * A0;
#pragma omp parallel for schedule(static)
    for (int i = HALF; i < X - HALF; i++) {
        for (int k = HALF; k < Z - HALF; k++) {
            const int j = k + i * Z;
            const REAL s1 = ((u[j - Z] + u[j + Z]) + u[j - 1]) + u[j + 1];
            const REAL s2 = ((u[j - 2 * Z] + u[j + 2 * Z]) + u[j - 2]) + u[j + 2];
            const REAL s3 = ((u[j - 3 * Z] + u[j + 3 * Z]) + u[j - 3]) + u[j + 3];
            const REAL s4 = ((u[j - 4 * Z] + u[j + 4 * Z]) + u[j - 4]) + u[j + 4];
            const REAL lap = (((c0 * u[j] + A1 * s1) + A2 * s2) + A3 * s3) + A4 * s4;
            U[j] = (C[j] * lap + (REAL)2.0f * u[j]) - U[j];
        }
    }
}

/* First-derivative-like difference, summed from m=4 down to m=1 as in
 * SRC/absorbingBoundaryCondition.c:183-186 (stride s picks the axis). */
static inline REAL FN(dfd)(const REAL *f, int j, int s)
{
    return ((B4 * (f[j + 4 * s] - f[j - 4 * s]) + B3 * (f[j + 3 * s] - f[j - 3 * s]))
            + B2 * (f[j + 2 * s] - f[j - 2 * s])) + B1 * (f[j + s] - f[j - s]);
}

/* One-axis second difference as in SRC/absorbingBoundaryCondition.c:208-212. */
static inline REAL FN(d2)(const REAL *f, int j, int s)
{
    return (((A4 * (f[j + 4 * s] + f[j - 4 * s]) + A3 * (f[j + 3 * s] + f[j - 3 * s]))
             + A2 * (f[j + 2 * s] + f[j - 2 * s])) + A1 * (f[j + s] + f[j - s])) + A0 * f[j];
}

/* CPML memory-variable update and corrections, in the order
 * Left, Top, Right, Bottom: SRC/absorbingBoundaryCondition.c:150-361.
 * u = waveField1 (u^n), U = waveField2 (post-stencil u^{n+1}). */
void FN(orc_cpml)(const orc_params *p, const REAL *u, REAL *U, const REAL *C,
                  const REAL *aL, const REAL *bL, const REAL *aR, const REAL *bR,
                  const REAL *aT, const REAL *bT, const REAL *aB, const REAL *bB,
                  REAL *psiL, REAL *psiR, REAL *psiT, REAL *psiB,
                  REAL *csiL, REAL *csiR, REAL *csiT, REAL *csiB)
{
    const int X = dimX(p), Z = dimZ(p);
    const int TL = p->t_left, TR = p->t_right, TU = p->t_up, TD = p->t_down;
    const REAL h = p->h;
    const REAL h1 = (REAL)1.0f / h, h2 = h * h, h3 = (REAL)1.0f / (h * h);
    const int XR = X - TR, ZD = Z - TD;

    /* Left: psi (:175-189), then dpsi/csi/U (:193-217) */
#pragma omp parallel for schedule(static)
    for (int i = HALF; i < TL - HALF; i++)
        for (int k = HALF; k < Z - HALF; k++) {
            const int q = Z * i + k;
            psiL[q] = bL[i] * psiL[q] + (aL[i] * FN(dfd)(u, k + i * Z, Z)) * h1;
        }
#pragma omp parallel for schedule(static)
    for (int i = HALF; i < TL - HALF; i++)
        for (int k = HALF; k < Z - HALF; k++) {
            const int q = Z * i + k, j = k + i * Z;
            const REAL dpsi = FN(dfd)(psiL, q, Z) * h1;
            csiL[q] = bL[i] * csiL[q] + aL[i] * (FN(d2)(u, j, Z) * h3 + dpsi);
            U[j] = U[j] + (C[j] * h2) * (dpsi + csiL[q]);
        }
    /* Top: psi for i in [4, X-4) (:221-235); corrections for i in [0, X) (:239-263) */
#pragma omp parallel for schedule(static)
    for (int i = HALF; i < X - HALF; i++)
        for (int k = HALF; k < TU - HALF; k++) {
            const int q = TU * i + k;
            psiT[q] = bT[k] * psiT[q] + (aT[k] * FN(dfd)(u, k + i * Z, 1)) * h1;
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < X; i++)
        for (int k = HALF; k < TU - HALF; k++) {
            const int q = TU * i + k, j = k + i * Z;
            const REAL dpsi = FN(dfd)(psiT, q, 1) * h1;
            csiT[q] = bT[k] * csiT[q] + aT[k] * (FN(d2)(u, j, 1) * h3 + dpsi);
            U[j] = U[j] + (C[j] * h2) * (dpsi + csiT[q]);
        }
    /* Right (:267-310) */
#pragma omp parallel for schedule(static)
    for (int i = X - (TR - HALF); i < X - HALF; i++)
        for (int k = HALF; k < Z - HALF; k++) {
            const int q = Z * (i - XR) + k;
            psiR[q] = bR[i - XR] * psiR[q] + (aR[i - XR] * FN(dfd)(u, k + i * Z, Z)) * h1;
        }
#pragma omp parallel for schedule(static)
    for (int i = X - (TR - HALF); i < X - HALF; i++)
        for (int k = HALF; k < Z - HALF; k++) {
            const int q = Z * (i - XR) + k, j = k + i * Z;
            const REAL dpsi = FN(dfd)(psiR, q, Z) * h1;
            csiR[q] = bR[i - XR] * csiR[q] + aR[i - XR] * (FN(d2)(u, j, Z) * h3 + dpsi);
            U[j] = U[j] + (C[j] * h2) * (dpsi + csiR[q]);
        }
    /* Bottom (:313-357) */
#pragma omp parallel for schedule(static)
    for (int i = HALF; i < X - HALF; i++)
        for (int k = Z - (TD - HALF); k < Z - HALF; k++) {
            const int q = TD * i + (k - ZD);
            psiB[q] = bB[k - ZD] * psiB[q] + (aB[k - ZD] * FN(dfd)(u, k + i * Z, 1)) * h1;
        }
#pragma omp parallel for schedule(static)
    for (int i = HALF; i < X - HALF; i++)
        for (int k = Z - (TD - HALF); k < Z - HALF; k++) {
            const int q = TD * i + (k - ZD), j = k + i * Z;
            const REAL dpsi = FN(dfd)(psiB, q, 1) * h1;
            csiB[q] = bB[k - ZD] * csiB[q] + aB[k - ZD] * (FN(d2)(u, j, 1) * h3 + dpsi);
            U[j] = U[j] + (C[j] * h2) * (dpsi + csiB[q]);
        }
}

/* I += S*R, Is += S*S over the unpadded region: SRC/image_conditions.c:20-33. */
void FN(orc_image)(const orc_params *p, REAL *I, REAL *Is, const REAL *S, const REAL *R)
{
    const int Z = dimZ(p);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < p->nx; i++)
        for (int k = 0; k < p->nz; k++) {
            const int j = (i + p->t_left) * Z + (k + p->t_up);
            I[i * p->nz + k] = I[i * p->nz + k] + S[j] * R[j];
            if (Is) Is[i * p->nz + k] = Is[i * p->nz + k] + S[j] * S[j];
        }
}

/* Kahan-compensated acc += v: the option p->image_kahan of this build (the
 * reference sums plainly).  The compensation c carries the low-order part
 * lost by the previous addition; every operation is one FP rounding
 * (-ffp-contract=off), in this order. */
static inline void FN(kahan_acc)(REAL *acc, REAL *c, REAL v)
{
    const REAL y = v - *c;
    const REAL t = *acc + y;
    *c = (t - *acc) - y;
    *acc = t;
}

/* Same, with the source wavefield given as an interior-only [Nx][Nz] array;
 * with compensations Ic, Isc (non-NULL) the sums are Kahan-compensated. */
static void FN(image_interior)(const orc_params *p, REAL *I, REAL *Is, const REAL *Sint, const REAL *R,
                               REAL *Ic, REAL *Isc)
{
    const int Z = dimZ(p);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < p->nx; i++)
        for (int k = 0; k < p->nz; k++) {
            const int j = (i + p->t_left) * Z + (k + p->t_up);
            const int q = i * p->nz + k;
            const REAL s = Sint[q];
            if (Ic) {
                FN(kahan_acc)(&I[q], &Ic[q], s * R[j]);
                if (Is) FN(kahan_acc)(&Is[q], &Isc[q], s * s);
            } else {
                I[q] = I[q] + s * R[j];
                if (Is) Is[q] = Is[q] + s * s;
            }
        }
}

typedef struct {
    REAL *C, *u, *U;
    REAL *aL, *aR, *aT, *aB, *bL, *bR, *bT, *bB;
    REAL *psiL, *psiR, *psiT, *psiB, *csiL, *csiR, *csiT, *csiB;
} FN(state);

static void FN(state_init)(FN(state) *s, const orc_params *p, const float *vpad)
{
    const int X = dimX(p), Z = dimZ(p);
    const size_t n = (size_t)X * Z;
    s->C = calloc(n, sizeof(REAL)); s->u = calloc(n, sizeof(REAL)); s->U = calloc(n, sizeof(REAL));
    s->aL = calloc(p->t_left, sizeof(REAL));  s->bL = calloc(p->t_left, sizeof(REAL));
    s->aR = calloc(p->t_right, sizeof(REAL)); s->bR = calloc(p->t_right, sizeof(REAL));
    s->aT = calloc(p->t_up, sizeof(REAL));    s->bT = calloc(p->t_up, sizeof(REAL));
    s->aB = calloc(p->t_down, sizeof(REAL));  s->bB = calloc(p->t_down, sizeof(REAL));
    s->psiL = calloc((size_t)p->t_left * Z, sizeof(REAL));  s->csiL = calloc((size_t)p->t_left * Z, sizeof(REAL));
    s->psiR = calloc((size_t)p->t_right * Z, sizeof(REAL)); s->csiR = calloc((size_t)p->t_right * Z, sizeof(REAL));
    s->psiT = calloc((size_t)X * p->t_up, sizeof(REAL));    s->csiT = calloc((size_t)X * p->t_up, sizeof(REAL));
    s->psiB = calloc((size_t)X * p->t_down, sizeof(REAL));  s->csiB = calloc((size_t)X * p->t_down, sizeof(REAL));
    FN(orc_coef)(p, s->C, vpad);
    FN(orc_profiles)(p, s->aL, s->aR, s->aT, s->aB, s->bL, s->bR, s->bT, s->bB);
}

static void FN(state_free)(FN(state) *s)
{
    REAL **all[] = {&s->C, &s->u, &s->U, &s->aL, &s->aR, &s->aT, &s->aB, &s->bL, &s->bR, &s->bT, &s->bB,
                    &s->psiL, &s->psiR, &s->psiT, &s->psiB, &s->csiL, &s->csiR, &s->csiT, &s->csiB};
    for (size_t t = 0; t < sizeof(all) / sizeof(all[0]); t++) { free(*all[t]); *all[t] = NULL; }
}

static void FN(state_step)(FN(state) *s, const orc_params *p)
{
    FN(orc_stencil)(p, s->C, s->u, s->U);
    FN(orc_cpml)(p, s->u, s->U, s->C, s->aL, s->bL, s->aR, s->bR, s->aT, s->bT, s->aB, s->bB,
                 s->psiL, s->psiR, s->psiT, s->psiB, s->csiL, s->csiR, s->csiT, s->csiB);
}

/* Forward (source) modeling: SRC/forward_modeling.c:98-130.
 * snaps (optional): interior-only snapshots [nt/TS][Nx][Nz] of U at n%TS==0.
 * seis  (optional): U at receiver r after the step, stored at seis[r*nt+n]
 *                   (the recording convention that mirrors the backward read
 *                   at SRC/backward_modeling.c:109).
 * last_fields (optional): u^{nt-1}, u^{nt} padded, for step-level checks. */
int FN(orc_forward)(const orc_params *p, int shot_x, int shot_z, const float *vpad,
                    const float *wavelet, REAL *snaps, REAL *seis, REAL *last_u, REAL *last_U)
{
    const int X = dimX(p), Z = dimZ(p), TS = p->time_slices;
    if (TS <= 0 || p->nt % TS != 0) return -1;
    FN(state) s;
    FN(state_init)(&s, p, vpad);
    const int xs = shot_x + p->t_left, zs = shot_z + p->t_up;
    const int zr = p->t_up + p->depth_receiver;
    const size_t nint = (size_t)p->nx * p->nz;
    for (int n = 0; n < p->nt; n++) {
        FN(state_step)(&s, p);
        if (n < p->n_wavelet)
            s.U[zs + xs * Z] = s.U[zs + xs * Z] + (REAL)wavelet[n] * s.C[xs * Z + zs];
        if (seis)
            for (int r = 0; r < p->n_receivers; r++) {
                const int pos = p->first_receiver + (p->spacing_receiver - 1) * r;
                seis[(size_t)r * p->nt + n] = s.U[(p->t_left + pos) * Z + zr];
            }
        if (snaps && n % TS == 0) {
            REAL *dst = snaps + (size_t)(n / TS) * nint;
            for (int i = 0; i < p->nx; i++)
                memcpy(dst + (size_t)i * p->nz, s.U + (size_t)(i + p->t_left) * Z + p->t_up, sizeof(REAL) * p->nz);
        }
        REAL *t = s.U; s.U = s.u; s.u = t;
    }
    if (last_u) memcpy(last_u, s.U, sizeof(REAL) * (size_t)X * Z);   /* after the final swap U holds u^{nt-1} */
    if (last_U) memcpy(last_U, s.u, sizeof(REAL) * (size_t)X * Z);
    FN(state_free)(&s);
    return 0;
}

/* Backward (receiver) modeling with imaging: SRC/backward_modeling.c:97-126.
 * Receiver injection is a replacement, not coef-scaled, at
 * pos_r = first + (spacing-1)*r (:106-110).  Imaging at tau%TS==0 pairs
 * snapshot tau/TS with the receiver field after injection (:112-120). */
int FN(orc_backward)(const orc_params *p, const float *vpad, const float *seis,
                     const REAL *snaps, REAL *I, REAL *Is, REAL *last_u, REAL *last_U)
{
    const int X = dimX(p), Z = dimZ(p), TS = p->time_slices;
    if (TS <= 0 || p->nt % TS != 0) return -1;
    FN(state) s;
    FN(state_init)(&s, p, vpad);
    const int zr = p->t_up + p->depth_receiver;
    const size_t nint = (size_t)p->nx * p->nz;
    /* Kahan option: compensations start at zero with each pass */
    REAL *Ic = p->image_kahan ? calloc(nint, sizeof(REAL)) : NULL;
    REAL *Isc = p->image_kahan ? calloc(nint, sizeof(REAL)) : NULL;
    if (p->image_kahan && (!Ic || !Isc)) {
        free(Ic); free(Isc); FN(state_free)(&s);
        return -2;
    }
    int c = p->nt / TS - 1;
    for (int n = 0; n < p->nt; n++) {
        const int tau = p->nt - n - 1;
        FN(state_step)(&s, p);
        for (int r = 0; r < p->n_receivers; r++) {
            const int pos = p->first_receiver + (p->spacing_receiver - 1) * r;
            s.U[(p->t_left + pos) * Z + zr] = (REAL)seis[(size_t)r * p->nt + tau];
        }
        if (tau % TS == 0) {
            if (snaps) FN(image_interior)(p, I, Is, snaps + (size_t)c * nint, s.U, Ic, Isc);
            c--;
        }
        REAL *t = s.U; s.U = s.u; s.u = t;
    }
    if (last_u) memcpy(last_u, s.U, sizeof(REAL) * (size_t)X * Z);
    if (last_U) memcpy(last_U, s.u, sizeof(REAL) * (size_t)X * Z);
    free(Ic);
    free(Isc);
    FN(state_free)(&s);
    return 0;
}

/* One shot of RTM: forward with in-memory snapshots, then backward+imaging.
 * I, Is are accumulated (+=), as adjointModeling does. */
int FN(orc_rtm_shot)(const orc_params *p, int shot_x, int shot_z, const float *vpad,
                     const float *wavelet, const float *seis, REAL *I, REAL *Is)
{
    const size_t nsnap = (size_t)(p->nt / (p->time_slices > 0 ? p->time_slices : 1));
    REAL *snaps = malloc(sizeof(REAL) * nsnap * (size_t)p->nx * p->nz);
    if (!snaps) return -2;
    int rc = FN(orc_forward)(p, shot_x, shot_z, vpad, wavelet, snaps, NULL, NULL, NULL);
    if (rc == 0) rc = FN(orc_backward)(p, vpad, seis, snaps, I, Is, NULL, NULL);
    free(snaps);
    return rc;
}

#if !defined(ORACLE_DOUBLE)
/* Edge-replication padding: SRC/generic_functions.c:5-48, including the
 * right-pad row range [T_U, Nz+T_D) of line 27. */
void orc_extend(const orc_params *p, float *ext, const float *model)
{
    const int Z = dimZ(p), TL = p->t_left, TU = p->t_up, TD = p->t_down, TR = p->t_right;
    const int Nx = p->nx, Nz = p->nz;
    for (int i = 0; i < Nx; i++)
        for (int k = 0; k < Nz; k++) ext[(k + TU) + (i + TL) * Z] = model[k + i * Nz];
    for (int i = 0; i < TL; i++)
        for (int k = TU; k < Nz + TU; k++) ext[k + i * Z] = ext[k + TL * Z];
    for (int i = Nx + TL; i < Nx + TL + TR; i++)
        for (int k = TU; k < Nz + TD; k++) ext[k + i * Z] = ext[k + (Nx + TL - 1) * Z];
    for (int i = 0; i < Nx + TL + TR; i++)
        for (int k = 0; k < TU; k++) ext[k + i * Z] = ext[TU + i * Z];
    for (int i = 0; i < Nx + TL + TR; i++)
        for (int k = Nz + TU; k < Nz + TU + TD; k++) ext[k + i * Z] = ext[(Nz + TU - 1) + i * Z];
}

/* 8th-order Laplacian post-filter of the stacked image (no coef, no h):
 * SRC/math_functions.c:9-27.  Border rows/columns are left untouched. */
void orc_laplacian(const orc_params *p, float *lp, const float *in)
{
    const int Nx = p->nx, Nz = p->nz;
    const float c0 = 2.0f * (-2.97399944f);
    for (int i = HALF; i < Nx - HALF; i++)
        for (int k = HALF; k < Nz - HALF; k++) {
            const int j = k + i * Nz;
            const float s1 = ((in[j - Nz] + in[j + Nz]) + in[j - 1]) + in[j + 1];
            const float s2 = ((in[j - 2 * Nz] + in[j + 2 * Nz]) + in[j - 2]) + in[j + 2];
            const float s3 = ((in[j - 3 * Nz] + in[j + 3 * Nz]) + in[j - 3]) + in[j + 3];
            const float s4 = ((in[j - 4 * Nz] + in[j + 4 * Nz]) + in[j - 4]) + in[j + 4];
            lp[j] = (((c0 * in[j] + (+1.70507669f) * s1) + (-0.25861812f) * s2) + (+0.04577745f) * s3)
                    + (-0.00523630f) * s4;
        }
}
#endif
