/*
 * rtm_main.c -- rtmuq2d_hip, the process driver: the reference's main
 * (SRC/mpi_rtm_iso2d.c:9-61) without MPI.  It runs rtm_routine(sim) for a range
 * of simulation ids, with the reference's split of ids over ranks
 * (:39-43), and can fan the ranks out over GPUs as child processes, one per
 * device.
 *
 *   rtmuq2d_hip [--params FILE] [--ranks N --rank R | --devices N] [--dry-run] [FIRST LAST]
 *
 * Parameters are the shipped SRC/fixed_parameters.h values, overridden by
 * FILE: one `name = value` per line ('#' starts a comment), name being an
 * rtm_params field (include/rtm_hip.h) or the reference's own macro name
 * (Nx, thicknessLeft, numberOfTimeStep, ...).  FIRST LAST override the id
 * range.  --dry-run prints the parameters and the ids each rank would run
 * and touches no GPU.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <sys/types.h>
#include <sys/wait.h>
#include <unistd.h>

#include "rtm_hip.h"

enum kind { K_INT, K_FLOAT, K_PATH };

static const struct field {
    const char *name, *macro;   /* rtm_params field, SRC/fixed_parameters.h macro */
    size_t off;
    enum kind kind;
} FIELDS[] = {
    {"velocity_file", "velocityFile", offsetof(rtm_params, velocity_file), K_PATH},
    {"wavelet_file", "waveletFile", offsetof(rtm_params, wavelet_file), K_PATH},
    {"seismogram_file", "seismogramFile", offsetof(rtm_params, seismogram_file), K_PATH},
    {"stacked_file", "stackedFile", offsetof(rtm_params, stacked_file), K_PATH},
    {"stacked_laplacian_file", "stackedLaplacianFile", offsetof(rtm_params, stacked_laplacian_file), K_PATH},
    {"zfp_tolerance", "zfp_tolerance", offsetof(rtm_params, zfp_tolerance), K_FLOAT},
    {"spacing_grid", "spacingGrid", offsetof(rtm_params, spacing_grid), K_FLOAT},
    {"sampling", "sampling", offsetof(rtm_params, sampling), K_FLOAT},
    {"number_of_simulations", "numberOfSimulations", offsetof(rtm_params, number_of_simulations), K_INT},
    {"nx", "Nx", offsetof(rtm_params, nx), K_INT},
    {"nz", "Nz", offsetof(rtm_params, nz), K_INT},
    {"first_location_shot", "firstLocationShot", offsetof(rtm_params, first_location_shot), K_INT},
    {"depth_shot", "depthShot", offsetof(rtm_params, depth_shot), K_INT},
    {"spacing_shot", "spacingShot", offsetof(rtm_params, spacing_shot), K_INT},
    {"number_of_shots", "numberOfShots", offsetof(rtm_params, number_of_shots), K_INT},
    {"first_location_receiver", "firstLocationReceiver", offsetof(rtm_params, first_location_receiver), K_INT},
    {"depth_receiver", "depthReceiver", offsetof(rtm_params, depth_receiver), K_INT},
    {"spacing_receiver", "spacingReceiver", offsetof(rtm_params, spacing_receiver), K_INT},
    {"number_of_receivers", "numberOfReceivers", offsetof(rtm_params, number_of_receivers), K_INT},
    {"number_of_time_steps", "numberOfTimeStep", offsetof(rtm_params, number_of_time_steps), K_INT},
    {"sampling_wavelet", "sampligWavelet", offsetof(rtm_params, sampling_wavelet), K_INT},
    {"cut_frequency", "cutFrequency", offsetof(rtm_params, cut_frequency), K_FLOAT},
    {"thickness_up", "thicknessUp", offsetof(rtm_params, thickness_up), K_INT},
    {"thickness_down", "thicknessDown", offsetof(rtm_params, thickness_down), K_INT},
    {"thickness_left", "thicknessLeft", offsetof(rtm_params, thickness_left), K_INT},
    {"thickness_right", "thicknessRight", offsetof(rtm_params, thickness_right), K_INT},
    {"time_slices", "TIME_SLICES", offsetof(rtm_params, time_slices), K_INT},
    {"snapshot_mode", NULL, offsetof(rtm_params, snapshot_mode), K_INT},
    {"image_kahan", NULL, offsetof(rtm_params, image_kahan), K_INT},
    {"device", NULL, offsetof(rtm_params, device), K_INT},
    {"normalize_illumination", NULL, offsetof(rtm_params, normalize_illumination), K_INT},
};
#define NFIELDS ((int)(sizeof FIELDS / sizeof FIELDS[0]))

static char *trim(char *s)
{
    while (*s == ' ' || *s == '\t') s++;
    char *e = s + strlen(s);
    while (e > s && (e[-1] == ' ' || e[-1] == '\t' || e[-1] == '\n' || e[-1] == '\r')) *--e = 0;
    return s;
}

/* apply a `name = value` file to p; 0 on success */
static int load_params(const char *path, rtm_params *p)
{
    FILE *f = fopen(path, "r");
    if (!f) { fprintf(stderr, "rtmuq2d_hip: cannot open %s\n", path); return -1; }
    char line[2 * RTM_PATH_LEN];
    int lineno = 0, rc = 0;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        char *s = trim(line);
        if (!*s) continue;
        char *eq = strchr(s, '=');
        if (!eq) { fprintf(stderr, "%s:%d: expected name = value\n", path, lineno); rc = -1; break; }
        *eq = 0;
        char *name = trim(s), *val = trim(eq + 1);
        const struct field *fd = NULL;
        for (int i = 0; i < NFIELDS && !fd; i++)
            if (!strcmp(name, FIELDS[i].name) || (FIELDS[i].macro && !strcmp(name, FIELDS[i].macro))) fd = &FIELDS[i];
        if (!fd) { fprintf(stderr, "%s:%d: unknown parameter '%s'\n", path, lineno, name); rc = -1; break; }
        char *end = NULL;
        void *dst = (char *)p + fd->off;
        if (fd->kind == K_INT) {
            const long v = strtol(val, &end, 10);
            if (end == val || *end) { fprintf(stderr, "%s:%d: '%s' is not an integer\n", path, lineno, val); rc = -1; break; }
            *(int *)dst = (int)v;
        } else if (fd->kind == K_FLOAT) {
            /* via double, as the reference's float globals are initialised
             * from double literals (SRC/fixed_parameters.h) */
            const float v = (float)strtod(val, &end);
            if (end == val || *end) { fprintf(stderr, "%s:%d: '%s' is not a number\n", path, lineno, val); rc = -1; break; }
            *(float *)dst = v;
        } else {
            size_t n = strlen(val);
            if (n >= 2 && val[0] == '"' && val[n - 1] == '"') { val[n - 1] = 0; val++; n -= 2; }
            if (n >= RTM_PATH_LEN) { fprintf(stderr, "%s:%d: path too long\n", path, lineno); rc = -1; break; }
            snprintf((char *)dst, RTM_PATH_LEN, "%s", val);
        }
    }
    fclose(f);
    return rc;
}

static void print_params(const rtm_params *p)
{
    for (int i = 0; i < NFIELDS; i++) {
        const void *src = (const char *)p + FIELDS[i].off;
        if (FIELDS[i].kind == K_INT) printf("%s = %d\n", FIELDS[i].name, *(const int *)src);
        else if (FIELDS[i].kind == K_FLOAT) printf("%s = %.9g\n", FIELDS[i].name, *(const float *)src);
        else printf("%s = %s\n", FIELDS[i].name, (const char *)src);
    }
}

/* the ids rank `rank` of `world` runs, SRC/mpi_rtm_iso2d.c:39-43 (float
 * arithmetic as there) */
static void rank_range(int nsim, int world, int rank, int *first, int *last)
{
    const float per = ((float)nsim) / world;
    *first = (int)(ceil((per * rank) + 1));
    *last = (int)(ceil(per * (rank + 1)));
}

static double now_s(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static int run_rank(const rtm_params *p, int rank, int first, int last)
{
    char host[256] = "localhost";
    (void)gethostname(host, sizeof host - 1);
    printf("[Host %s] [rank %d] %s \n", host, rank, "STARTING 2D ISOTROPIC REVERSE TIME MIGRATION !!!");
    fflush(stdout);
    if (rtm_set_params(p)) { fprintf(stderr, "rtmuq2d_hip: %s\n", rtm_last_error()); return 1; }
    const double t0 = now_s();
    for (int sim = first; sim <= last; sim++) rtm_routine(sim);
    printf("--> The process %d took %lf seconds to run.\n", rank, now_s() - t0);
    fflush(stdout);
    return 0;
}

static void usage(void)
{
    fprintf(stderr, "usage: rtmuq2d_hip [--params FILE] [--ranks N --rank R | --devices N] [--dry-run] [FIRST LAST]\n");
}

int main(int argc, char **argv)
{
    rtm_params p;
    rtm_params_default(&p);
    int ranks = 1, rank = 0, devices = 0, dry = 0, first = -1, last = -1, npos = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "--params") && i + 1 < argc) {
            if (load_params(argv[++i], &p)) return 2;
        } else if (!strcmp(a, "--ranks") && i + 1 < argc) {
            ranks = atoi(argv[++i]);
        } else if (!strcmp(a, "--rank") && i + 1 < argc) {
            rank = atoi(argv[++i]);
        } else if (!strcmp(a, "--devices") && i + 1 < argc) {
            devices = atoi(argv[++i]);
        } else if (!strcmp(a, "--dry-run")) {
            dry = 1;
        } else if (a[0] != '-' && npos < 2) {
            if (npos++ == 0) first = atoi(a);
            else last = atoi(a);
        } else {
            usage();
            return 2;
        }
    }
    if (npos == 1 || ranks < 1 || rank < 0 || rank >= ranks || devices < 0 || (devices && ranks > 1)) {
        usage();
        return 2;
    }
    const int nsim = npos ? last - first + 1 : p.number_of_simulations;
    const int base = npos ? first - 1 : 0;
    const int world = devices ? devices : ranks;
    if (nsim < 1) { fprintf(stderr, "rtmuq2d_hip: empty simulation range\n"); return 2; }
    if (dry) {
        print_params(&p);
        for (int r = (devices ? 0 : rank); r < (devices ? world : rank + 1); r++) {
            int f, l;
            rank_range(nsim, world, r, &f, &l);
            printf("rank %d device %d simulations %d %d\n", r, devices ? r : p.device, base + f, base + l);
        }
        return 0;
    }
    if (!devices) {
        int f, l;
        rank_range(nsim, ranks, rank, &f, &l);
        return run_rank(&p, rank, base + f, base + l);
    }
    /* one child per device, started before this process touches any GPU */
    fflush(stdout);
    pid_t *pids = calloc((size_t)devices, sizeof(pid_t));
    if (!pids) return 1;
    for (int d = 0; d < devices; d++) {
        pids[d] = fork();
        if (pids[d] < 0) { perror("rtmuq2d_hip: fork"); return 1; }
        if (pids[d] == 0) {
            int f, l;
            rank_range(nsim, devices, d, &f, &l);
            p.device = d;
            _exit(run_rank(&p, d, base + f, base + l));
        }
    }
    int bad = 0;
    for (int d = 0; d < devices; d++) {
        int st = 0;
        if (waitpid(pids[d], &st, 0) < 0 || !WIFEXITED(st) || WEXITSTATUS(st) != 0) {
            fprintf(stderr, "rtmuq2d_hip: device %d process failed\n", d);
            bad = 1;
        }
    }
    free(pids);
    return bad;
}
