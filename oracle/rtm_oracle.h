/*
 * rtm_oracle.h -- CPU restatement of rtmUq2D's RTM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is
 * compared against; it is never linked into, called by, or shipped with the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Layout follows the reference exactly: padded fields are [X][Z] with z
 * fastest (index k + i*Z, SRC/stencilDFM.c:23), CPML strips use the
 * reference's own strip layouts (SRC/absorbingBoundaryCondition.c:180,226,
 * 272,319) and images are [Nx][Nz] z fastest (SRC/image_conditions.c:25).
 *
 * Two builds of the same source exist: float (orc_*_f32, the reference's
 * arithmetic) and double (orc_*_f64, the same scheme in FP64).  Both are
 * compiled with -ffp-contract=off so every expression is evaluated exactly
 * in the order the reference writes it.
 */
#ifndef RTM_ORACLE_H
#define RTM_ORACLE_H

#include <stdint.h>

typedef struct orc_params {
    int nx, nz;                 /* model size (Nx, Nz)            fixed_parameters.h:31-32 */
    int t_left, t_right;        /* CPML thickness                 fixed_parameters.h:67-70 */
    int t_up, t_down;
    float h, dt;                /* spacingGrid, sampling          fixed_parameters.h:23-24 */
    int nt;                     /* numberOfTimeStep               fixed_parameters.h:53    */
    int n_wavelet;              /* sampligWavelet                 fixed_parameters.h:57    */
    int first_receiver;         /* fixed_parameters.h:45-48 */
    int depth_receiver;
    int spacing_receiver;
    int n_receivers;
    int time_slices;            /* TIME_SLICES                    forward_modeling.h:4     */
    int image_kahan;            /* 0: I += S*R as the reference (image_conditions.c:29-30);
                                   1: the same sums Kahan-compensated (an option of this
                                   build, not of the reference) */
} orc_params;

#endif
