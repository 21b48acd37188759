/*
 * rtm_hip.h -- C ABI of librtm_hip.so, the MI355X (gfx950) drop-in for the
 * reverse-time-migration hot path of rtmUq2D.
 *
 * Two layers are exported, both with plain C types (no HIP or torch types):
 *
 *  1. The reference's own entry points, unchanged in name, signature, memory
 *     ownership and observable behaviour.  Each cites the declaration it
 *     replaces.  Sizes come from the process-wide rtm_params (set with
 *     rtm_set_params; the defaults are SRC/fixed_parameters.h:8-70), just as
 *     the reference reads them from its compile-time globals.
 *
 *  2. The device-resident API (rtm_plan_*, rtm_batch_*): wavefields, CPML
 *     state and the source-snapshot store live on HBM for a whole
 *     forward/backward pass; only inputs and images cross PCIe.  These have no
 *     counterpart in the reference and return an int status (0 = success,
 *     negative = error; the message is in rtm_last_error()).
 *
 * Layouts at this boundary are the reference's: padded fields [X][Z] with z
 * fastest (SRC/stencilDFM.c:23), models and images [Nx][Nz] z fastest
 * (SRC/image_conditions.c:25), seismograms [nrec][nt] t fastest
 * (SRC/backward_modeling.c:109).  Device layouts are internal.
 */
#ifndef RTM_HIP_H
#define RTM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTM_PATH_LEN 512

/* Snapshot store precision (stepping is always FP32 in an FP32 plan). */
enum rtm_snapshot_mode {
    RTM_SNAP_FP32 = 0,      /* snapshot as computed                        */
    RTM_SNAP_FP16 = 1,      /* FP16 with a per-snapshot max-abs scale      */
    RTM_SNAP_BF16 = 2,      /* BF16 with a per-snapshot max-abs scale      */
    RTM_SNAP_BOUNDARY = 3,  /* boundary saving: 4-point ring, reconstruct  */
    RTM_SNAP_I16 = 4        /* int16 with a power-of-two scale per 64-point
                               run of an image row (tile shapes only)     */
};

enum rtm_real { RTM_REAL_F32 = 0, RTM_REAL_F64 = 1 };

/* Runtime form of SRC/fixed_parameters.h:8-70 plus TIME_SLICES
 * (SRC/forward_modeling.h:4).  rtm_params_default() fills the reference's
 * shipped values. */
typedef struct rtm_params {
    char velocity_file[RTM_PATH_LEN];        /* fixed_parameters.h:8  (printf pattern, sim id) */
    char wavelet_file[RTM_PATH_LEN];         /* fixed_parameters.h:9  */
    char seismogram_file[RTM_PATH_LEN];      /* fixed_parameters.h:10 (printf pattern, shot id) */
    char stacked_file[RTM_PATH_LEN];         /* fixed_parameters.h:14 */
    char stacked_laplacian_file[RTM_PATH_LEN]; /* fixed_parameters.h:15 */
    float zfp_tolerance;                     /* fixed_parameters.h:19 */
    float spacing_grid;                      /* h,  fixed_parameters.h:23 */
    float sampling;                          /* dt, fixed_parameters.h:24 */
    int number_of_simulations;               /* fixed_parameters.h:28 */
    int nx, nz;                              /* fixed_parameters.h:31-32 */
    int first_location_shot, depth_shot, spacing_shot, number_of_shots;           /* :37-40 */
    int first_location_receiver, depth_receiver, spacing_receiver, number_of_receivers; /* :45-48 */
    int number_of_time_steps;                /* fixed_parameters.h:53 */
    int sampling_wavelet;                    /* fixed_parameters.h:57 */
    float cut_frequency;                     /* fixed_parameters.h:58 (the CPML uses its own 60 Hz) */
    int thickness_up, thickness_down, thickness_left, thickness_right; /* :67-70 */
    int time_slices;                         /* TIME_SLICES, forward_modeling.h:4 */
    /* device-side options, no reference counterpart */
    int snapshot_mode;                       /* enum rtm_snapshot_mode */
    int image_kahan;                         /* 1: Kahan-compensated FP32 image accumulation */
    int device;                              /* HIP device ordinal */
    /* rtm_routine: 1 divides the stacked image by the stacked self-correlation
     * (the reference's full-survey line, commented out at SRC/rtm_iso2d.c:150;
     * points the source never illuminated stay 0); 0 writes the raw stack as
     * the shipped reference does */
    int normalize_illumination;
} rtm_params;

void rtm_params_default(rtm_params *p);
int rtm_set_params(const rtm_params *p);      /* process-wide params for the legacy symbols */
int rtm_get_params(rtm_params *p);
const char *rtm_last_error(void);             /* thread-local message of the last failure */
int rtm_device_count(void);

/* ---------------------------------------------------------------------------
 * 1. Reference entry points (same names and signatures)
 * ------------------------------------------------------------------------- */

/* SRC/rtm_iso2d.h:11 -- one simulation: read inputs, shot loop, write images. */
void rtm_routine(int simulationId);

/* SRC/forward_modeling.h:11 -- source modeling; the snapshot store that the
 * reference writes to TEMPORARY_FILES/sourceWavefield_%05d.bin is kept on HBM,
 * keyed by simId, until adjointModeling(simId, ...) consumes it. */
void isotropicAcousticModeling(int simId, int *shot, float *velocity, float *wavelet);

/* SRC/backward_modeling.h:11 -- receiver modeling + imaging; images are
 * accumulated (+=) into the caller's arrays; frees simId's snapshot store.
 * Returns 1, as the reference does (SRC/backward_modeling.c:135,164). */
int adjointModeling(int simId, int *shot, float *velocity, float *wavelet, float *seismogram,
                    float *imageCondition, float *selfImageCondition);

/* SRC/stencilDFM.h:15 -- one interior leapfrog step, host pointers (H2D, step, D2H). */
void iso_acoustic_wave_equation(float *pond, float *waveField2, float *waveField3);

/* SRC/absorbingBoundaryCondition.h:27 -- CPML damping profiles (host). */
void attenuation_profile_cpml(float *a_Left, float *a_Right, float *a_Top, float *a_Bottom,
                              float *b_Left, float *b_Right, float *b_Top, float *b_Bottom);

/* SRC/absorbingBoundaryCondition.h:29 -- one CPML update, host pointers (H2D, step, D2H). */
void aiso_attenuation_cpml(float *waveField1, float *waveField2, float *Coef,
                           float *a_left, float *b_left, float *a_right, float *b_right,
                           float *a_top, float *b_top, float *a_bottom, float *b_bottom,
                           float *psi_borderLeft, float *psi_borderRight, float *psi_borderTop,
                           float *psi_borderBottom, float *csi_borderLeft, float *csi_borderRight,
                           float *csi_borderTop, float *csi_borderBottom);

/* SRC/image_conditions.h:4-6 -- imaging conditions, host pointers. */
void cross_correlation(float *imageCondition, float *sourceWavefield, float *receiverWavefield);
void normalized_cross_correlation(float *imageCondition, float *selfImageCondition,
                                  float *sourceWavefield, float *receiverWavefield);

/* SRC/compression_tools.h:6,9 -- headerless ZFP 2-D fixed-accuracy streams
 * (host; the ZFP 0.5.5 stream format, written and read by this library's own
 * codec below, without the ZFP library).  Decompression returns 0 on
 * success, 1 on failure (also when the file ends before the field does). */
int zfp2d_lossyCompression(void **compression, float *array, int dimension1, int dimension2,
                           float tolerance);
int zfp2d_lossyDecompression(char *zfp_data_02, float *array, int dimension1, int dimension2,
                             float tolerance);

/* The codec behind them, in memory (no reference counterpart: the reference
 * calls the ZFP library, ZFP_LIBRARY_0.5.5/include/zfp.h).  Streams of
 * a[x + nx*y] as zfp_field_2d(a, zfp_type_float, nx, ny) with 64-bit stream
 * words; the three lossy modes of zfp_stream_set_accuracy / _precision /
 * _rate (rate without write-random-access alignment). */
typedef struct rtm_zfp_mode {
    unsigned minbits, maxbits, maxprec;
    int minexp;
} rtm_zfp_mode;
double rtm_zfp_mode_accuracy(rtm_zfp_mode *m, double tol);       /* returns the tolerance used */
unsigned rtm_zfp_mode_precision(rtm_zfp_mode *m, unsigned prec);  /* returns the precision used */
double rtm_zfp_mode_rate(rtm_zfp_mode *m, double rate);           /* returns the rate used */
/* zfp_stream_maximum_size for such a field (bytes) */
size_t rtm_zfp2f_max_bytes(const rtm_zfp_mode *m, unsigned nx, unsigned ny);
/* stream bytes (a multiple of 8), or 0 if cap is too small; any thread
 * count gives the same bytes */
size_t rtm_zfp2f_encode(const rtm_zfp_mode *m, const float *a, unsigned nx, unsigned ny, void *out, size_t cap,
                        int threads);
/* bytes consumed (whole words), or 0 if the stream ends before the field */
size_t rtm_zfp2f_decode(const rtm_zfp_mode *m, const void *in, size_t bytes, unsigned nx, unsigned ny, float *a);

/* SRC/pi_file.h:5,8 -- raw float32 files; exit(1) on any error, as the reference. */
int reading_file(char *path, float *matrix, int dim1, int dim2);
int recording_file(char *path, float *matrix, int dim1, int dim2);

/* SRC/generic_functions.h:5 -- pad the model with CPML strips by edge replication. */
void extend_matrix(float *extendedModel, float *model);

/* SRC/math_functions.h:15 -- 8th-order Laplacian post-filter of the image. */
void laplacian_filter2D(float *lp, float *inputMatrix);

/* ---------------------------------------------------------------------------
 * 2. Device-resident API (no reference counterpart)
 * ------------------------------------------------------------------------- */

typedef struct rtm_plan rtm_plan;

/* One simulation's device state: wavefields, coef, CPML memory variables,
 * wavelet, seismogram, snapshot store and image.  real = enum rtm_real. */
int rtm_plan_create(const rtm_params *p, int real, rtm_plan **out);
int rtm_plan_destroy(rtm_plan *plan);
size_t rtm_plan_footprint(const rtm_params *p, int real);   /* HBM bytes a plan allocates */

/* Which kernel steps the interior (the points outside every CPML strip):
 * TILES   the LDS tile kernel, as everywhere else;
 * COLUMNS wavefronts march down 128-column strips with a register window
 *         and lane shifts (no LDS, no barrier), beside the strip tiles;
 * MIXED   one tile launch with 64 x 4*rows_per_thread interior tiles and
 *         64 x 4 tiles wherever a CPML strip is touched;
 * TILES8  tiles of 64 x 8*rows_per_thread in workgroups of 8 wavefronts;
 * PAIRS   the whole tiles of the image region (but those within 4 points of
 *         its edge) are stepped two time steps per launch by wavefronts
 *         marching 112-column strips (segments of seg_rows rows), reading
 *         u^n, u^{n-1} and C once for both steps; every other tile one step
 *         per launch.  Four wavefield buffers per simulation; FP32 fields and
 *         the FP32 snapshot store only (the request server does not take it). */
enum rtm_interior {
    RTM_INTERIOR_TILES = 0, RTM_INTERIOR_COLUMNS = 1, RTM_INTERIOR_MIXED = 2, RTM_INTERIOR_TILES8 = 3,
    RTM_INTERIOR_PAIRS = 4
};

/* Kernel shape: rows_per_thread in {1,2,4,8} gives tiles of 64 x
 * 4*rows_per_thread points (MIXED: 2 or 4, for the interior tiles);
 * graph_block = time steps per captured hipGraph block (<= 1: one eager
 * launch per step); interior = enum rtm_interior; seg_rows = rows per column
 * segment (COLUMNS, PAIRS; <= 0: whole strips).  Tiles stay wherever a CPML strip is
 * touched.  Defaults: 4, 100, TILES, 0.  Results are bit-identical for every
 * shape. */
int rtm_plan_set_tiling(rtm_plan *plan, int rows_per_thread, int graph_block, int interior, int seg_rows);
int rtm_plan_get_tiling(rtm_plan *plan, int *rows_per_thread, int *graph_block, int *interior, int *seg_rows);
/* Autotuner: time every candidate shape (steps forward time steps from rest,
 * no source and no stores, at the plan's graph block) on this plan's grid,
 * keep the fastest, and remember it for plans of the same grid class created
 * later in the process.  shapes [cap][4] = {rows_per_thread, graph_block,
 * interior, seg_rows} and ms_per_step [cap] of the candidates (either may be
 * NULL), *n_out = entries written.  Leaves the wavefields (not the image or
 * snapshots) overwritten. */
int rtm_plan_autotune(rtm_plan *plan, int steps, int cap, int *shapes, double *ms_per_step, int *n_out);

int rtm_plan_set_velocity(rtm_plan *plan, const float *vpad);     /* padded [X][Z] */
int rtm_plan_set_wavelet(rtm_plan *plan, const float *w, int n);
int rtm_plan_set_seismogram(rtm_plan *plan, const float *d);      /* [nrec][nt] */
int rtm_plan_zero_image(rtm_plan *plan);

/* nt forward steps from rest with the source at model (shot_x, shot_z);
 * stores snapshots; if record != 0 also records the seismogram on device. */
int rtm_plan_forward(rtm_plan *plan, int shot_x, int shot_z, int record);
/* nt backward steps from rest with receiver replacement and imaging;
 * accumulates into the device image. */
int rtm_plan_backward(rtm_plan *plan);
/* forward + backward for one shot */
int rtm_plan_migrate(rtm_plan *plan, int shot_x, int shot_z);

int rtm_plan_get_seismogram(rtm_plan *plan, float *d);             /* [nrec][nt] */
int rtm_plan_get_image(rtm_plan *plan, float *I, float *Is);      /* [Nx][Nz]; Is may be NULL */
int rtm_plan_get_image_f64(rtm_plan *plan, double *I, double *Is);
int rtm_plan_get_fields(rtm_plan *plan, float *u_prev, float *u_last);  /* padded [X][Z] */
int rtm_plan_get_fields_f64(rtm_plan *plan, double *u_prev, double *u_last);
int rtm_plan_synchronize(rtm_plan *plan);
/* Shot stacking on the device (replaces the per-shot temporary image files
 * and the host stack of SRC/rtm_iso2d.c:115-146): rtm_plan_stack_image adds
 * the plan's current images (I, I_s) into its stacks, in call order, which
 * is the reference's shot order; the stacks start at zero.
 * rtm_plan_get_stack returns the stack, divided by the self-image stack
 * where that is non-zero and 0 elsewhere if normalize (SRC/rtm_iso2d.c:150),
 * and its 8th-order Laplacian (laplacian_filter2D, SRC/math_functions.c:9-27),
 * both [Nx][Nz]; either output may be NULL. */
int rtm_plan_stack_image(rtm_plan *plan);
int rtm_plan_zero_stack(rtm_plan *plan);
int rtm_plan_get_stack(rtm_plan *plan, int normalize, float *stack, float *laplacian);
/* Average time per time step (ms) of the last forward and the last backward
 * pass, from HIP events recorded on the plan's stream around the step
 * launches; steps = time steps per pass. */
int rtm_plan_last_timing(rtm_plan *plan, double *fwd_step_ms, double *bwd_step_ms, int *steps);

/* Ensemble of nsims realizations of one geometry (the reference's
 * simulations, SRC/mpi_rtm_iso2d.c:39-51, here on one device): every
 * realization has its own velocity, wavefields, CPML state, snapshot store
 * and image; the shot gather (wavelet, traces) is shared.  The realizations
 * of a launch group advance together, one tiled launch per time step in
 * hipGraph blocks. */
typedef struct rtm_batch rtm_batch;

int rtm_batch_create(const rtm_params *p, int nsims, int real, rtm_batch **out);
int rtm_batch_destroy(rtm_batch *batch);
size_t rtm_batch_footprint(const rtm_params *p, int nsims, int real);
int rtm_batch_size(rtm_batch *batch);
int rtm_batch_set_tiling(rtm_batch *batch, int rows_per_thread, int graph_block, int interior, int seg_rows);
int rtm_batch_get_tiling(rtm_batch *batch, int *rows_per_thread, int *graph_block, int *interior, int *seg_rows);
/* The autotuner over the realizations of one launch group (as
 * rtm_plan_autotune; ms_per_step is the time per step of that group). */
int rtm_batch_autotune(rtm_batch *batch, int steps, int cap, int *shapes, double *ms_per_step, int *n_out);
/* Launch groups (no reference counterpart): the realizations run in groups
 * of at most sims_per_group (0 or >= the batch size: one group), as few
 * groups as that allows with sizes that differ by at most one (64
 * realizations at sims_per_group 15: 13, 13, 13, 13, 12), each group's
 * forward and backward passes before the next group's.  get_group reports
 * the largest group's size and the number of groups.  Results do not
 * depend on the grouping.  A group whose per-step working set fits the
 * device's last-level cache steps faster than one that does not.  A new
 * batch starts with the group size its grids were last tuned to, else with
 * as many realizations as fit 224 MiB at 16 B (FP32) or 32 B (FP64) per
 * padded grid point. */
int rtm_batch_set_group(rtm_batch *batch, int sims_per_group);
int rtm_batch_get_group(rtm_batch *batch, int *sims_per_group, int *groups);
/* Shape-tunes each candidate group size (1/2 to 2 times the size that fills
 * the 224 MiB budget, each taken to the largest group of its balanced split,
 * and the whole batch) on group 0, timing forward steps that write
 * the snapshot stores (overwritten by the next migration); each group within
 * 10% of the fastest is timed a second time at its tuned shape and the
 * slower reading counts.  Keeps the smallest group within 2% of the least
 * time per realization-step; row k = group_sizes[k], us_per_sim_step[k].
 * It runs one group at a time (lanes 1). */
int rtm_batch_autotune_group(rtm_batch *batch, int steps, int cap, int *group_sizes, double *us_per_sim_step,
                             int *n_out);
/* Launch groups in flight at once (no reference counterpart): with lanes > 1
 * (at most 4) group k runs on lane k % lanes, each lane a stream of its own,
 * so one group's steps fill the gaps between another's step launches; the
 * group count becomes a multiple of the lanes.  A migration then reports the
 * span of all its passes, half as forward and half as backward time
 * (rtm_batch_last_timing).  Results do not depend on the lanes. */
int rtm_batch_set_lanes(rtm_batch *batch, int lanes);
int rtm_batch_get_lanes(rtm_batch *batch, int *lanes);
/* Times one lane at the current group size, then two lanes at each distinct
 * balanced group size from 1/8 to 5/8 of the size that fills the 224 MiB
 * budget and three lanes from 1/8 to 1/4 (at most about 13 sizes per lane
 * count), each as a wave of its first groups in flight, on the host clock
 * in two sweeps over the candidates (the slowest reading counts), at
 * the current shape, stepping forward steps that write the snapshot stores;
 * keeps one lane if it is within 2% of the least time per realization-step,
 * else the fastest.  Row k = lanes[k], group_sizes[k] (the largest group),
 * us_per_sim_step[k]. */
int rtm_batch_autotune_lanes(rtm_batch *batch, int steps, int cap, int *lanes, int *group_sizes,
                             double *us_per_sim_step, int *n_out);
int rtm_batch_set_velocity(rtm_batch *batch, int sim, const float *vpad);   /* padded [X][Z] */
int rtm_batch_set_wavelet(rtm_batch *batch, const float *w, int n);
int rtm_batch_set_seismogram(rtm_batch *batch, const float *d);             /* [nrec][nt] */
int rtm_batch_zero_images(rtm_batch *batch);
int rtm_batch_migrate(rtm_batch *batch, int shot_x, int shot_z);
int rtm_batch_get_image(rtm_batch *batch, int sim, float *I, float *Is);     /* [Nx][Nz] */
int rtm_batch_get_image_f64(rtm_batch *batch, int sim, double *I, double *Is);
/* mean and unbiased variance of the realizations' images, FP64 Welford in
 * realization order, [Nx][Nz] */
int rtm_batch_welford(rtm_batch *batch, double *mean, double *var);
/* rtm_batch_welford with no device allocation: no reference counterpart;
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same results, bit for bit.  Its
 * partial results live in a fixed device array of the library (32 MiB +
 * 32 KiB per device, held until the process ends): at most 4096
 * realizations and 1,048,576 image points (Nx*Nz), larger batches are
 * refused with an error.  All calls on one device share that array, so
 * calls that overlap on one device race. */
int rtm_batch_welford_noalloc(rtm_batch *batch, double *mean, double *var);
int rtm_batch_last_timing(rtm_batch *batch, double *fwd_step_ms, double *bwd_step_ms, int *steps);

/* ---------------------------------------------------------------------------
 * 3. Request server (no reference counterpart): continuous batching of
 *    migrations of any grid size on one device.
 * ------------------------------------------------------------------------- */

typedef struct rtm_server rtm_server;

/* A server owning an HBM arena of hbm_bytes and max_slots request slots;
 * every request must use snapshot_mode (enum rtm_snapshot_mode). */
int rtm_server_create(int device, size_t hbm_bytes, int max_slots, int snapshot_mode, rtm_server **out);
int rtm_server_destroy(rtm_server *server);
/* Kernel shape shared by all requests (as rtm_plan_set_tiling); graph_block
 * is also the number of time steps per tick. */
int rtm_server_set_tiling(rtm_server *server, int rows_per_thread, int graph_block, int interior, int seg_rows);
/* Arena bytes a request with params p occupies while active. */
size_t rtm_server_request_footprint(const rtm_params *p);
/* Queue one migration: params (grid, time, receivers; snapshot_mode as the
 * server's), padded velocity [X][Z], wavelet [sampling_wavelet], seismogram
 * [nrec][nt] and the shot position.  The inputs are copied.  Returns the
 * request id (>= 0), or < 0 when the request can never run (the source or a
 * receiver outside the stepped region, or larger than the arena). */
int rtm_server_submit(rtm_server *server, const rtm_params *p, const float *vpad, const float *wavelet,
                      const float *seismogram, int shot_x, int shot_z);
/* Admit queued requests that fit (arrival order), advance every active
 * request by one block of time steps in shared launches, switch finished
 * forward passes to their backward pass and retire finished requests.
 * Returns the number of requests still queued or active (< 0 only for a
 * null server).  Failures retire the requests they touch, not the server:
 * a device error in one launch group's step fails that group's requests; a
 * failed upload, or a request that does not fit even the empty arena, fails
 * that request; each is reported by rtm_server_take_image, and the other
 * requests and later ticks carry on. */
int rtm_server_tick(rtm_server *server);
/* 1 and *id = the oldest retired (completed or failed) request not yet
 * taken, or 0 if none. */
int rtm_server_next_done(rtm_server *server, int *id);
/* Copy a completed request's image ([Nx][Nz], n = Nx*Nz floats) and forget
 * it: 0; -2 (and forgotten) if the request failed, the reason in
 * rtm_last_error(); -1 if it is not retired or n differs. */
int rtm_server_take_image(rtm_server *server, int id, float *image, size_t n);
int rtm_server_stats(rtm_server *server, int *active, int *queued, size_t *bytes_in_use, long long *steps);
/* Requests failed since the server was created. */
int rtm_server_failures(rtm_server *server);
/* Threads: rtm_server_submit, _next_done and _take_image may be called from
 * any thread at any time, concurrently with a tick (a submission copies its
 * inputs and queues them; it never waits for a tick's launches).  Ticks,
 * stats and the configuration calls serialise on one lock.  Destroy only
 * once no other thread calls into the server. */
/* Launch groups: the active requests advance in groups whose per-step
 * working sets (16 B per padded grid point) sum to at most `bytes`, group
 * after group, each a block of steps; a group holds requests of one pass
 * (forward or backward), so its launches have one pass's store or imaging
 * steps.  0 puts every active request in one group.  The default, 224 MiB,
 * keeps a group within the device's 256 MB last-level cache.  Results do
 * not depend on the grouping. */
int rtm_server_set_group_bytes(rtm_server *server, size_t bytes);
int rtm_server_get_group_bytes(rtm_server *server, size_t *bytes, int *groups);
/* Launch groups in flight at once (1 to 4; default 1): group g of a tick
 * steps on lane g % lanes, a stream of its own, so the tick's groups run
 * together.  Results do not depend on the lanes. */
int rtm_server_set_lanes(rtm_server *server, int lanes);
int rtm_server_get_lanes(rtm_server *server, int *lanes);

#ifdef __cplusplus
}
#endif

#endif /* RTM_HIP_H */
