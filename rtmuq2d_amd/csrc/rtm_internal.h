/*
 * rtm_internal.h -- declarations shared by the host C layer (rtm_host.c) and
 * the HIP device layer (rtm_device.hip).  Not part of the public ABI.
 */
#ifndef RTM_INTERNAL_H
#define RTM_INTERNAL_H

#include "rtm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* error message of the calling thread (printf-style) */
void rtm_set_error(const char *fmt, ...);

/* validated params: non-zero with an error set when unusable */
int rtm_params_check(const rtm_params *p);

/* CPML profiles of SRC/absorbingBoundaryCondition.c:11-148 for params p, in
 * the reference's per-side arrays (length = thickness of that side).  The
 * f64 form evaluates the same formulas in double precision. */
void rtm_profiles_f32(const rtm_params *p, float *aL, float *aR, float *aT, float *aB,
                      float *bL, float *bR, float *bT, float *bB);
void rtm_profiles_f64(const rtm_params *p, double *aL, double *aR, double *aT, double *aB,
                      double *bL, double *bR, double *bT, double *bB);

/* process-wide params behind the legacy symbols */
const rtm_params *rtm_global_params(void);

/* legacy per-step symbols, device side (host pointers in reference layout) */
int rtmd_stencil_host(const rtm_params *p, const float *pond, const float *u, float *U);
int rtmd_cpml_host(const rtm_params *p, const float *u, float *U, const float *C,
                   const float *aL, const float *bL, const float *aR, const float *bR,
                   const float *aT, const float *bT, const float *aB, const float *bB,
                   float *psiL, float *psiR, float *psiT, float *psiB,
                   float *csiL, float *csiR, float *csiT, float *csiB);
int rtmd_image_host(const rtm_params *p, float *I, float *Is, const float *S, const float *R);

/* snapshot-store table keyed by simulation id (replaces the
 * TEMPORARY_FILES/sourceWavefield_%05d.bin hand-off) */
int rtmd_table_put(int sim_id, rtm_plan *plan);   /* replaces and frees any previous plan */
rtm_plan *rtmd_table_take(int sim_id);            /* removes and returns, or NULL */

#ifdef __cplusplus
}
#endif

#endif
