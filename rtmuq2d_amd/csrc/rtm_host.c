/*
 * rtm_host.c -- host C layer of librtm_hip.so.
 *
 * Exports rtmUq2D's own entry points with their names, signatures and
 * observable behaviour, and runs their arithmetic through the device layer
 * (rtm_device.hip).  Host-kept pieces stay here, in C: the runtime parameter
 * block that replaces the compile-time globals of SRC/fixed_parameters.h, the
 * CPML profiles, model padding, raw float32 file I/O, the ZFP 0.5.5 seismogram
 * codec and the per-simulation driver rtm_routine.
 */
#define _GNU_SOURCE
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rtm_internal.h"
#include "rtm_zfp.h"

/* ------------------------------------------------------------------------ */
/* parameters                                                                */
/* ------------------------------------------------------------------------ */

/* The reference's shipped configuration, SRC/fixed_parameters.h:8-70, plus
 * TIME_SLICES = 20 (SRC/forward_modeling.h:4). */
void rtm_params_default(rtm_params *p)
{
    memset(p, 0, sizeof *p);
    snprintf(p->velocity_file, RTM_PATH_LEN, "%s", "./INPUTS/vp_marmousi2_%05d.bin");
    snprintf(p->wavelet_file, RTM_PATH_LEN, "%s", "./INPUTS/i_ricker.rsf@");
    snprintf(p->seismogram_file, RTM_PATH_LEN, "%s", "./INPUTS/seismogram_tol10-0_%05d.zfp");
    snprintf(p->stacked_file, RTM_PATH_LEN, "%s", "./OUTPUTS/stackedBrute_%05d.bin");
    snprintf(p->stacked_laplacian_file, RTM_PATH_LEN, "%s", "./OUTPUTS/stackedLaplacian_%05d.bin");
    p->zfp_tolerance = 0.0f;
    p->spacing_grid = 6.25f;
    p->sampling = 0.0005f;
    p->number_of_simulations = 2;
    p->nx = 2721;
    p->nz = 561;
    p->first_location_shot = 5;
    p->depth_shot = 5;
    p->spacing_shot = 17;
    p->number_of_shots = 1;
    p->first_location_receiver = 1;
    p->depth_receiver = 5;
    p->spacing_receiver = 3;
    p->number_of_receivers = 1361;
    p->number_of_time_steps = 14000;
    p->sampling_wavelet = 800;
    p->cut_frequency = 50.0f;
    p->thickness_up = p->thickness_down = p->thickness_left = p->thickness_right = 50;
    p->time_slices = 20;
    p->snapshot_mode = RTM_SNAP_FP32;
    p->image_kahan = 0;
    p->device = 0;
    p->normalize_illumination = 0;
}

static rtm_params g_params;
static int g_params_init = 0;

const rtm_params *rtm_global_params(void)
{
    if (!g_params_init) {
        rtm_params_default(&g_params);
        g_params_init = 1;
    }
    return &g_params;
}

int rtm_params_check(const rtm_params *p)
{
    if (!p) { rtm_set_error("params: null pointer"); return -1; }
    if (p->nx < 9 || p->nz < 9) { rtm_set_error("params: nx=%d nz=%d (need >= 9)", p->nx, p->nz); return -1; }
    /* the CPML needs at least one active strip point per side:
     * active range [4, T-4), SRC/absorbingBoundaryCondition.c:75 */
    if (p->thickness_left < 9 || p->thickness_right < 9 || p->thickness_up < 9 || p->thickness_down < 9) {
        rtm_set_error("params: CPML thickness must be >= 9 on every side (got L%d R%d U%d D%d)",
                      p->thickness_left, p->thickness_right, p->thickness_up, p->thickness_down);
        return -1;
    }
    if (!(p->spacing_grid > 0.0f) || !(p->sampling > 0.0f)) {
        rtm_set_error("params: spacing_grid and sampling must be > 0");
        return -1;
    }
    if (p->number_of_time_steps <= 0 || p->time_slices <= 0 ||
        p->number_of_time_steps % p->time_slices != 0) {
        /* with nt % TS != 0 the reference pairs snapshots off by one
         * (SRC/backward_modeling.c:96 vs :112); refuse that case */
        rtm_set_error("params: number_of_time_steps (%d) must be a positive multiple of time_slices (%d)",
                      p->number_of_time_steps, p->time_slices);
        return -1;
    }
    if (p->sampling_wavelet < 0 || p->number_of_receivers < 0) {
        rtm_set_error("params: negative sampling_wavelet or number_of_receivers");
        return -1;
    }
    if (p->snapshot_mode < RTM_SNAP_FP32 || p->snapshot_mode > RTM_SNAP_I16) {
        rtm_set_error("params: unknown snapshot_mode %d", p->snapshot_mode);
        return -1;
    }
    return 0;
}

int rtm_set_params(const rtm_params *p)
{
    if (rtm_params_check(p)) return -1;
    g_params = *p;
    g_params_init = 1;
    return 0;
}

int rtm_get_params(rtm_params *p)
{
    if (!p) { rtm_set_error("rtm_get_params: null pointer"); return -1; }
    *p = *rtm_global_params();
    return 0;
}

static void die_on(int rc, const char *who)
{
    if (rc) {
        printf("\nError in %s: %s\n", who, rtm_last_error());
        exit(1);
    }
}

/* ------------------------------------------------------------------------ */
/* CPML damping profiles, SRC/absorbingBoundaryCondition.c:11-148           */
/* ------------------------------------------------------------------------ */
/* d is the normalised distance into the layer (1 at the outer edge).
 * sigma = smax d^2, alpha = amax (1 - d) + amax/10, b = e^{-(sigma+alpha) dt},
 * a = sigma (b - 1) / (sigma + alpha).  smax = 3 pi (N+1) vmax / L_left for
 * every side (line 65), vmax = 4000, fcut = 60 (lines 42-43, which shadow the
 * global cutFrequency), f_dom = fcut / (3 sqrt(pi)). */

static void side_f32(float *a, float *b, int T, int outer_first, float h, float dt, float smax, float amax)
{
    const float Lm = (float)(T - 8) * h;
    for (int t = 0; t < T; t++) { a[t] = 0.0f; b[t] = 0.0f; }
    for (int t = 0; t < T - 8; t++) {
        float d;
        if (outer_first) d = (Lm - h * (float)t) / Lm;       /* left / top: index t+4      */
        else d = (h + h * (float)t) / Lm;                    /* right / bottom: index t+4  */
        const float sg = smax * powf(d, 2.0f);
        const float al = amax * (1.0f - d) + 0.1f * amax;
        const float bb = powf((float)M_E, -(sg + al) * dt);
        b[t + 4] = bb;
        a[t + 4] = sg * (bb - 1.0f) / (sg + al);
    }
}

void rtm_profiles_f32(const rtm_params *p, float *aL, float *aR, float *aT, float *aB,
                      float *bL, float *bR, float *bT, float *bB)
{
    const float pi = (float)M_PI;
    const float h = p->spacing_grid, dt = p->sampling;
    const float fdom = 60.0f / (3.0f * sqrt(pi));
    const float Lleft = (float)(p->thickness_left - 8) * h;
    const float smax = (3.0f * pi) * (((2.0f + 1.0f) * 4000.0f) / Lleft);
    const float amax = pi * fdom;
    side_f32(aL, bL, p->thickness_left, 1, h, dt, smax, amax);
    side_f32(aR, bR, p->thickness_right, 0, h, dt, smax, amax);
    side_f32(aT, bT, p->thickness_up, 1, h, dt, smax, amax);
    side_f32(aB, bB, p->thickness_down, 0, h, dt, smax, amax);
}

static void side_f64(double *a, double *b, int T, int outer_first, double h, double dt, double smax,
                     double amax)
{
    const double Lm = (double)(T - 8) * h;
    for (int t = 0; t < T; t++) { a[t] = 0.0; b[t] = 0.0; }
    for (int t = 0; t < T - 8; t++) {
        const double d = outer_first ? (Lm - h * (double)t) / Lm : (h + h * (double)t) / Lm;
        const double sg = smax * pow(d, 2.0);
        const double al = amax * (1.0 - d) + (double)0.1f * amax;
        const double bb = exp(-(sg + al) * dt);
        b[t + 4] = bb;
        a[t + 4] = sg * (bb - 1.0) / (sg + al);
    }
}

void rtm_profiles_f64(const rtm_params *p, double *aL, double *aR, double *aT, double *aB,
                      double *bL, double *bR, double *bT, double *bB)
{
    const double pi = M_PI;
    const double h = p->spacing_grid, dt = p->sampling;
    const double fdom = 60.0 / (3.0 * sqrt(pi));
    const double Lleft = (double)(p->thickness_left - 8) * h;
    const double smax = (3.0 * pi) * ((3.0 * 4000.0) / Lleft);
    const double amax = pi * fdom;
    side_f64(aL, bL, p->thickness_left, 1, h, dt, smax, amax);
    side_f64(aR, bR, p->thickness_right, 0, h, dt, smax, amax);
    side_f64(aT, bT, p->thickness_up, 1, h, dt, smax, amax);
    side_f64(aB, bB, p->thickness_down, 0, h, dt, smax, amax);
}

void attenuation_profile_cpml(float *a_Left, float *a_Right, float *a_Top, float *a_Bottom,
                              float *b_Left, float *b_Right, float *b_Top, float *b_Bottom)
{
    rtm_profiles_f32(rtm_global_params(), a_Left, a_Right, a_Top, a_Bottom, b_Left, b_Right, b_Top,
                     b_Bottom);
}

/* ------------------------------------------------------------------------ */
/* legacy per-step kernels (host pointers, device arithmetic)                */
/* ------------------------------------------------------------------------ */

void iso_acoustic_wave_equation(float *pond, float *waveField2, float *waveField3)
{
    die_on(rtmd_stencil_host(rtm_global_params(), pond, waveField2, waveField3),
           "iso_acoustic_wave_equation");
}

void aiso_attenuation_cpml(float *waveField1, float *waveField2, float *Coef,
                           float *a_left, float *b_left, float *a_right, float *b_right,
                           float *a_top, float *b_top, float *a_bottom, float *b_bottom,
                           float *psi_borderLeft, float *psi_borderRight, float *psi_borderTop,
                           float *psi_borderBottom, float *csi_borderLeft, float *csi_borderRight,
                           float *csi_borderTop, float *csi_borderBottom)
{
    die_on(rtmd_cpml_host(rtm_global_params(), waveField1, waveField2, Coef, a_left, b_left, a_right,
                          b_right, a_top, b_top, a_bottom, b_bottom, psi_borderLeft, psi_borderRight,
                          psi_borderTop, psi_borderBottom, csi_borderLeft, csi_borderRight,
                          csi_borderTop, csi_borderBottom),
           "aiso_attenuation_cpml");
}

void normalized_cross_correlation(float *imageCondition, float *selfImageCondition,
                                  float *sourceWavefield, float *receiverWavefield)
{
    die_on(rtmd_image_host(rtm_global_params(), imageCondition, selfImageCondition, sourceWavefield,
                           receiverWavefield),
           "normalized_cross_correlation");
}

void cross_correlation(float *imageCondition, float *sourceWavefield, float *receiverWavefield)
{
    die_on(rtmd_image_host(rtm_global_params(), imageCondition, NULL, sourceWavefield, receiverWavefield),
           "cross_correlation");
}

/* ------------------------------------------------------------------------ */
/* modeling entry points                                                     */
/* ------------------------------------------------------------------------ */

void isotropicAcousticModeling(int simId, int *shot, float *velocity, float *wavelet)
{
    const rtm_params *p = rtm_global_params();
    rtm_plan *plan = NULL;
    die_on(rtm_plan_create(p, RTM_REAL_F32, &plan), "isotropicAcousticModeling");
    int rc = rtm_plan_set_velocity(plan, velocity);
    if (rc == 0) rc = rtm_plan_set_wavelet(plan, wavelet, p->sampling_wavelet);
    if (rc == 0) rc = rtm_plan_forward(plan, shot[0], shot[1], 0);
    if (rc) {
        rtm_plan_destroy(plan);
        die_on(rc, "isotropicAcousticModeling");
    }
    rtmd_table_put(simId, plan);   /* the HBM snapshot store, keyed like the temp file */
}

int adjointModeling(int simId, int *shot, float *velocity, float *wavelet, float *seismogram,
                    float *imageCondition, float *selfImageCondition)
{
    (void)shot;     /* unused by the reference's backward pass as well */
    (void)wavelet;  /* SRC/backward_modeling.c: the wavelet argument is unused */
    const rtm_params *p = rtm_global_params();
    rtm_plan *plan = rtmd_table_take(simId);
    if (!plan) {
        printf("\nError: no source wavefield stored for simulation %d\n", simId);
        return 1;
    }
    const size_t n = (size_t)p->nx * p->nz;
    int rc = rtm_plan_set_velocity(plan, velocity);
    if (rc == 0) rc = rtm_plan_set_seismogram(plan, seismogram);
    if (rc == 0) rc = rtm_plan_zero_image(plan);
    if (rc == 0) rc = rtm_plan_backward(plan);
    float *I = NULL, *Is = NULL;
    if (rc == 0) {
        I = malloc(sizeof(float) * n);
        Is = malloc(sizeof(float) * n);
        if (!I || !Is) { rtm_set_error("adjointModeling: out of host memory"); rc = -1; }
    }
    if (rc == 0) rc = rtm_plan_get_image(plan, I, Is);
    if (rc == 0)
        for (size_t t = 0; t < n; t++) {   /* accumulate (+=), as the reference */
            imageCondition[t] = imageCondition[t] + I[t];
            selfImageCondition[t] = selfImageCondition[t] + Is[t];
        }
    free(I);
    free(Is);
    rtm_plan_destroy(plan);   /* the reference removes the snapshot file (:131) */
    die_on(rc, "adjointModeling");
    return 1;
}

/* ------------------------------------------------------------------------ */
/* files, padding, post-filter                                               */
/* ------------------------------------------------------------------------ */

/* raw float32 read of dim1*dim2 values; exit(1) on any failure (SRC/pi_file.c:11-35) */
int reading_file(char *path, float *matrix, int dim1, int dim2)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        printf("\nError oppening file (fopen in reading_file function)\n");
        exit(1);
    }
    const size_t want = (size_t)dim1 * (size_t)dim2;
    const size_t got = fread(matrix, sizeof(float), want, f);
    fclose(f);
    if (got != want) {
        printf("\nError reading file (fread in reading_file function)\n");
        exit(1);
    }
    return 0;
}

/* raw float32 write; exit(1) on any failure (SRC/pi_file.c:37-61) */
int recording_file(char *path, float *matrix, int dim1, int dim2)
{
    FILE *f = fopen(path, "wb");
    if (!f) {
        printf("\nError oppening file (fopen in recording_file function)\n");
        exit(1);
    }
    const size_t want = (size_t)dim1 * (size_t)dim2;
    const size_t put = fwrite(matrix, sizeof(float), want, f);
    const int closed = fclose(f);
    if (put != want || closed != 0) {
        printf("\nError writing file (fwrite in recording_file function)\n");
        exit(1);
    }
    return 0;
}

/* Pad [Nx][Nz] to [X][Z] by edge replication, SRC/generic_functions.c:5-48:
 * interior copy; left and right strips copy the edge column over model rows;
 * then top and bottom strips copy the edge row over the full width.  The
 * right strip covers rows [T_U, Nz+T_D) as the reference's loop does (:27). */
static void extend_with(const rtm_params *p, float *ext, const float *model)
{
    const int Nx = p->nx, Nz = p->nz, TL = p->thickness_left, TR = p->thickness_right,
              TU = p->thickness_up, TD = p->thickness_down;
    const int Z = TU + Nz + TD, X = TL + Nx + TR;
    for (int i = 0; i < Nx; i++)
        memcpy(ext + (size_t)(i + TL) * Z + TU, model + (size_t)i * Nz, sizeof(float) * Nz);
    for (int i = 0; i < TL; i++)
        for (int k = TU; k < TU + Nz; k++) ext[(size_t)i * Z + k] = ext[(size_t)TL * Z + k];
    for (int i = TL + Nx; i < X; i++)
        for (int k = TU; k < Nz + TD; k++) ext[(size_t)i * Z + k] = ext[(size_t)(TL + Nx - 1) * Z + k];
    for (int i = 0; i < X; i++) {
        for (int k = 0; k < TU; k++) ext[(size_t)i * Z + k] = ext[(size_t)i * Z + TU];
        for (int k = TU + Nz; k < Z; k++) ext[(size_t)i * Z + k] = ext[(size_t)i * Z + TU + Nz - 1];
    }
}

void extend_matrix(float *extendedModel, float *model)
{
    extend_with(rtm_global_params(), extendedModel, model);
}

/* 8th-order Laplacian of an [Nx][Nz] image over [4,Nx-4) x [4,Nz-4), no coef
 * and no h (SRC/math_functions.c:9-27); other entries are not written. */
static void laplacian_with(const rtm_params *p, float *lp, const float *in)
{
    const int Nx = p->nx, Nz = p->nz;
    const float w0 = 2.0f * -2.97399944f, w1 = 1.70507669f, w2 = -0.25861812f, w3 = 0.04577745f,
                w4 = -0.00523630f;
#pragma omp parallel for schedule(static)
    for (int i = 4; i < Nx - 4; i++) {
        const float *c = in + (size_t)i * Nz;
        for (int k = 4; k < Nz - 4; k++) {
            const float t1 = ((c[k - Nz] + c[k + Nz]) + c[k - 1]) + c[k + 1];
            const float t2 = ((c[k - 2 * Nz] + c[k + 2 * Nz]) + c[k - 2]) + c[k + 2];
            const float t3 = ((c[k - 3 * Nz] + c[k + 3 * Nz]) + c[k - 3]) + c[k + 3];
            const float t4 = ((c[k - 4 * Nz] + c[k + 4 * Nz]) + c[k - 4]) + c[k + 4];
            lp[(size_t)i * Nz + k] = (((w0 * c[k] + w1 * t1) + w2 * t2) + w3 * t3) + w4 * t4;
        }
    }
}

void laplacian_filter2D(float *lp, float *inputMatrix)
{
    laplacian_with(rtm_global_params(), lp, inputMatrix);
}

/* ------------------------------------------------------------------------ */
/* ZFP 0.5.5 streams (host): this library's codec, rtm_zfp.c                 */
/* ------------------------------------------------------------------------ */

static int zfp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Headerless fixed-accuracy stream of the 2-D field (nx = dimension1 fastest,
 * ny = dimension2) over the flat array, as SRC/compression_tools.c:10-74 asks
 * ZFP for (OpenMP execution: rows of blocks are encoded in parallel and their
 * bits concatenated, the serial encoder's bytes).  Returns the stream size;
 * *compression receives a malloc'ed buffer of the maximum stream size. */
int zfp2d_lossyCompression(void **compression, float *array, int dimension1, int dimension2,
                           float tolerance)
{
    *compression = NULL;
    if (dimension1 <= 0 || dimension2 <= 0) return 0;
    rtm_zfp_mode m;
    rtm_zfp_mode_accuracy(&m, tolerance);
    const size_t cap = rtm_zfp2f_max_bytes(&m, (unsigned)dimension1, (unsigned)dimension2);
    void *buf = malloc(cap);
    size_t size = 0;
    if (buf) size = rtm_zfp2f_encode(&m, array, (unsigned)dimension1, (unsigned)dimension2, buf, cap, zfp_threads());
    *compression = buf;
    return (int)size;
}

/* Decode the headerless stream in `zfp_data_02` into array (same field
 * convention and tolerance as the encoder; the mode is not stored in the
 * stream), SRC/compression_tools.c:78-135: at most the maximum stream size
 * is read.  0 = success, 1 = failure ("Decompression Failed" on stderr, as
 * the reference). */
int zfp2d_lossyDecompression(char *zfp_data_02, float *array, int dimension1, int dimension2,
                             float tolerance)
{
    int status = 1;
    if (dimension1 > 0 && dimension2 > 0) {
        rtm_zfp_mode m;
        rtm_zfp_mode_accuracy(&m, tolerance);
        const size_t cap = rtm_zfp2f_max_bytes(&m, (unsigned)dimension1, (unsigned)dimension2);
        void *buf = malloc(cap ? cap : 1);
        FILE *f = buf ? fopen(zfp_data_02, "rb") : NULL;
        if (f) {
            const size_t got = fread(buf, 1, cap, f);
            fclose(f);
            status = rtm_zfp2f_decode(&m, buf, got, (unsigned)dimension1, (unsigned)dimension2, array) ? 0 : 1;
        }
        free(buf);
    }
    if (status) fprintf(stderr, "Decompression Failed\n");
    return status;
}

/* ------------------------------------------------------------------------ */
/* per-simulation driver, SRC/rtm_iso2d.c:15-177                             */
/* ------------------------------------------------------------------------ */

/* One simulation: read velocity and wavelet, pad, then for each shot run the
 * forward pass, decode the shot's seismogram, run the backward pass and stack
 * the shot image; write the raw stack and its Laplacian.  The device plan is
 * kept for all shots; the shot images are stacked on the device instead of
 * going through the reference's temporary files, summed in shot order as the
 * reference does (SRC/rtm_iso2d.c:115-129), and the illumination division
 * and the Laplacian run there too. */
void rtm_routine(int simulationId)
{
    const rtm_params *p = rtm_global_params();
    const int X = p->thickness_left + p->nx + p->thickness_right;
    const int Z = p->thickness_up + p->nz + p->thickness_down;
    const size_t nimg = (size_t)p->nx * p->nz;
    char path[RTM_PATH_LEN + 64];

    float *model = calloc(nimg, sizeof(float));
    float *vpad = calloc((size_t)X * Z, sizeof(float));
    float *seis = calloc((size_t)p->number_of_receivers * p->number_of_time_steps, sizeof(float));
    float *wav = calloc((size_t)(p->sampling_wavelet > 0 ? p->sampling_wavelet : 1), sizeof(float));
    float *stack = calloc(nimg, sizeof(float)), *lap = calloc(nimg, sizeof(float));
    if (!model || !vpad || !seis || !wav || !stack || !lap) {
        printf("\nError: out of host memory in rtm_routine\n");
        exit(1);
    }
    printf("\n\n.................................... STARTING 2D ISOTROPIC REVERSE TIME MIGRATION !!! ..........\n");
    snprintf(path, sizeof path, p->velocity_file, simulationId);
    reading_file(path, model, p->nx, p->nz);
    snprintf(path, sizeof path, "%s", p->wavelet_file);
    reading_file(path, wav, 1, p->sampling_wavelet);
    extend_with(p, vpad, model);

    rtm_plan *plan = NULL;
    die_on(rtm_plan_create(p, RTM_REAL_F32, &plan), "rtm_routine");
    die_on(rtm_plan_set_velocity(plan, vpad), "rtm_routine");
    die_on(rtm_plan_set_wavelet(plan, wav, p->sampling_wavelet), "rtm_routine");
    for (int shot = 1; shot <= p->number_of_shots; shot++) {
        printf("SHOT %d\n", shot);
        const int sx = p->first_location_shot + (shot - 1) * p->spacing_shot;
        die_on(rtm_plan_forward(plan, sx, p->depth_shot, 0), "rtm_routine");
        snprintf(path, sizeof path, p->seismogram_file, shot);
        (void)zfp2d_lossyDecompression(path, seis, p->number_of_receivers, p->number_of_time_steps,
                                       p->zfp_tolerance);
        die_on(rtm_plan_set_seismogram(plan, seis), "rtm_routine");
        die_on(rtm_plan_zero_image(plan), "rtm_routine");
        die_on(rtm_plan_backward(plan), "rtm_routine");
        die_on(rtm_plan_stack_image(plan), "rtm_routine");
    }
    printf("\n\n");
    /* SRC/rtm_iso2d.c:150 with unlit points left at 0, then SRC/rtm_iso2d.c:155-165 */
    die_on(rtm_plan_get_stack(plan, p->normalize_illumination, stack, lap), "rtm_routine");
    rtm_plan_destroy(plan);

    snprintf(path, sizeof path, p->stacked_file, simulationId);
    recording_file(path, stack, p->nx, p->nz);
    snprintf(path, sizeof path, p->stacked_laplacian_file, simulationId);
    recording_file(path, lap, p->nx, p->nz);

    free(model); free(vpad); free(seis); free(wav); free(stack); free(lap);
}
