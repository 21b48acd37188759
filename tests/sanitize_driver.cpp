// sanitize_driver.cpp -- runs the device kernel bodies (through the CPU
// harness, tests/cpu_step_harness.cpp) and the host C layer
// (rtmuq2d_amd/csrc/rtm_host.c, rtm_zfp.c) on the CPU so that they can be
// built with -fsanitize=address,undefined: every LDS index, halo read, psi
// halo and clamped global index of the tile, column and mixed-tile
// phases, on grids that exercise every tile class and partial tiles.  Test
// infrastructure only (tests/test_sanitizers.py builds and runs it); it
// checks memory and UB, not values (the parity tests check values).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include <unistd.h>

#include "cpu_step_harness.cpp"

extern "C" {
#include "rtm_internal.h"

// The device layer is not linked here: its entry points that rtm_host.c
// references fail, which the pure-host functions exercised below never reach.
void rtm_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    va_end(ap);
}
const char *rtm_last_error(void) { return ""; }
int rtm_plan_create(const rtm_params *, int, rtm_plan **out) { *out = nullptr; return -1; }
int rtm_plan_destroy(rtm_plan *) { return 0; }
int rtm_plan_forward(rtm_plan *, int, int, int) { return -1; }
int rtm_plan_backward(rtm_plan *) { return -1; }
int rtm_plan_get_image(rtm_plan *, float *, float *) { return -1; }
int rtm_plan_get_stack(rtm_plan *, int, float *, float *) { return -1; }
int rtm_plan_set_seismogram(rtm_plan *, const float *) { return -1; }
int rtm_plan_set_velocity(rtm_plan *, const float *) { return -1; }
int rtm_plan_set_wavelet(rtm_plan *, const float *, int) { return -1; }
int rtm_plan_stack_image(rtm_plan *) { return -1; }
int rtm_plan_zero_image(rtm_plan *) { return -1; }
int rtmd_stencil_host(const rtm_params *, const float *, const float *, float *) { return -1; }
int rtmd_cpml_host(const rtm_params *, const float *, float *, const float *, const float *, const float *,
                   const float *, const float *, const float *, const float *, const float *, const float *, float *,
                   float *, float *, float *, float *, float *, float *, float *)
{
    return -1;
}
int rtmd_image_host(const rtm_params *, float *, float *, const float *, const float *) { return -1; }
int rtmd_table_put(int, rtm_plan *) { return -1; }
rtm_plan *rtmd_table_take(int) { return nullptr; }
}

namespace {

std::mt19937 rng(5);

template <typename R>
void fill(std::vector<R> &v, double scale)
{
    std::normal_distribution<double> d(0.0, scale);
    for (auto &x : v) x = R(d(rng));
}

struct Grid {
    int nx, nz, TL, TR, TU, TD;
};

// one configuration's device-layout buffers, random state
template <typename R>
struct Bufs {
    int X, Z, P, nt, TS, nrec, rec_row;
    std::vector<R> f[22];     // the harness's pointer table order (see run_*)
    std::vector<int> recmap;
    std::vector<float> wav, seis_t, seis_rec;
    std::vector<unsigned short> q16;
    std::vector<unsigned> smax;
    Bufs(const Grid &g, int nt_, int TS_)
    {
        X = g.TL + g.nx + g.TR; Z = g.TU + g.nz + g.TD; P = (X + 63) / 64 * 64;
        nt = nt_; TS = TS_;
        nrec = g.nx / 3;
        rec_row = g.TU + 2;
        const size_t nf = (size_t)P * Z, nint = (size_t)g.nx * g.nz;
        for (int t = 0; t < 22; t++) f[t].assign(nf, R(0));
        f[17].assign(nint * (nt / TS), R(0));          // snapshots
        for (int t = 18; t < 22; t++) f[t].assign(nint, R(0));
        for (auto &v : f[4]) v = R(0.05);             // C
        for (int t : {5, 6}) { f[t].assign(X, R(0)); fill(f[t], 0.5); }   // ax, bx
        for (int t : {7, 8}) { f[t].assign(Z, R(0)); fill(f[t], 0.5); }   // az, bz
        recmap.assign(X, -1);
        for (int r = 0; r < nrec; r++) recmap[g.TL + 2 * r] = r;
        wav.assign(nt, 0.f);
        for (auto &w : wav) w = 1.0f;
        seis_t.assign((size_t)nt * nrec, 0.5f);
        seis_rec.assign((size_t)nt * nrec, 0.f);
        q16.assign(nint * (nt / TS), 0);
        smax.assign((size_t)(nt / TS) * rtm::kSnapWords, 0);
    }
};

template <typename R>
void run_tiles_grid(const Grid &g, int rz, int interior, int seg, int qmode, bool f64)
{
    const int nt = 6, TS = 2;
    Bufs<R> B(g, nt, TS);
    fill(B.f[0], 1.0);
    fill(B.f[1], 1.0);
    const R hs[3] = {R(0.1), R(100), R(0.01)};
    // table of hx_tiles: u0 u1 C ax bx az bz pXi pXo pZi pZo cX cZ snaps I Is Ic Isc q16 smax
    R *ptr[20] = {B.f[0].data(), B.f[1].data(), B.f[4].data(), B.f[5].data(), B.f[6].data(), B.f[7].data(),
                  B.f[8].data(), B.f[9].data(), B.f[10].data(), B.f[11].data(), B.f[12].data(), B.f[13].data(),
                  B.f[14].data(), B.f[17].data(), B.f[18].data(), B.f[19].data(), B.f[20].data(), B.f[21].data(),
                  (R *)(void *)B.q16.data(), (R *)(void *)B.smax.data()};
    const int src_j = (g.TU + 3) * B.P + g.TL + g.nx / 2;
    for (int phase = 1; phase <= 2; phase++)
        for (int n = 0; n < nt; n++) {
            const int flags = phase == 1 ? (4 | 2) : (8 | 1);
            int gg[22] = {B.X, B.Z, B.P, g.TL, g.TR, g.TU, g.TD, g.nx, g.nz, nt, TS, nt, B.nrec, phase, flags,
                          src_j, B.rec_row, n, rz, seg, interior, qmode};
            if (f64)
                hx_tiles_f64(gg, (const double *)(const void *)hs, (double *const *)(void *)ptr, B.recmap.data(),
                             B.wav.data(), B.seis_t.data(), B.seis_rec.data());
            else
                hx_tiles_f32(gg, (const float *)(const void *)hs, (float *const *)(void *)ptr, B.recmap.data(),
                             B.wav.data(), B.seis_t.data(), B.seis_rec.data());
        }
}

void run_host(const Grid &g)
{
    rtm_params p;
    rtm_params_default(&p);
    p.nx = g.nx; p.nz = g.nz;
    p.thickness_left = g.TL; p.thickness_right = g.TR; p.thickness_up = g.TU; p.thickness_down = g.TD;
    p.number_of_receivers = g.nx / 2; p.first_location_receiver = 0; p.spacing_receiver = 2;
    p.number_of_time_steps = 20; p.sampling_wavelet = 20; p.time_slices = 4;
    if (rtm_set_params(&p)) { fprintf(stderr, "params refused\n"); exit(2); }
    std::vector<float> aL(g.TL), aR(g.TR), aT(g.TU), aB(g.TD), bL(g.TL), bR(g.TR), bT(g.TU), bB(g.TD);
    rtm_profiles_f32(&p, aL.data(), aR.data(), aT.data(), aB.data(), bL.data(), bR.data(), bT.data(), bB.data());
    std::vector<double> dL(g.TL), dR(g.TR), dT(g.TU), dB(g.TD), eL(g.TL), eR(g.TR), eT(g.TU), eB(g.TD);
    rtm_profiles_f64(&p, dL.data(), dR.data(), dT.data(), dB.data(), eL.data(), eR.data(), eT.data(), eB.data());
    attenuation_profile_cpml(aL.data(), aR.data(), aT.data(), aB.data(), bL.data(), bR.data(), bT.data(), bB.data());
    const int X = g.TL + g.nx + g.TR, Z = g.TU + g.nz + g.TD;
    std::vector<float> model((size_t)g.nx * g.nz), ext((size_t)X * Z), lap((size_t)g.nx * g.nz);
    fill(model, 1.0);
    extend_matrix(ext.data(), model.data());
    laplacian_filter2D(lap.data(), model.data());
    char path[] = "/tmp/rtm_sanitize_XXXXXX";
    const int fd = mkstemp(path);
    if (fd >= 0) {
        recording_file(path, model.data(), g.nx, g.nz);
        reading_file(path, lap.data(), g.nx, g.nz);
        // the seismogram codec through the reference's call convention
        void *z = nullptr;
        const int n = zfp2d_lossyCompression(&z, model.data(), g.nx, g.nz, 1e-3f);
        FILE *f = fopen(path, "wb");
        if (f && z) { fwrite(z, 1, (size_t)n, f); fclose(f); }
        free(z);
        zfp2d_lossyDecompression(path, lap.data(), g.nx, g.nz, 1e-3f);
        unlink(path);
    }
    // the codec in every mode, ragged shapes, serial and parallel
    for (int mode = 0; mode < 3; mode++) {
        rtm_zfp_mode m;
        if (mode == 0) rtm_zfp_mode_accuracy(&m, 1e-4);
        else if (mode == 1) rtm_zfp_mode_precision(&m, 9);
        else rtm_zfp_mode_rate(&m, 5.5);
        const size_t cap = rtm_zfp2f_max_bytes(&m, g.nx, g.nz);
        std::vector<unsigned char> buf(cap);
        for (int threads : {1, 4}) {
            const size_t bytes = rtm_zfp2f_encode(&m, model.data(), g.nx, g.nz, buf.data(), cap, threads);
            rtm_zfp2f_decode(&m, buf.data(), bytes, g.nx, g.nz, lap.data());
            rtm_zfp2f_decode(&m, buf.data(), bytes / 2, g.nx, g.nz, lap.data());   // a short stream
        }
    }
}

}  // namespace

int main()
{
    const Grid grids[] = {{64, 48, 16, 16, 16, 16}, {72, 40, 14, 20, 18, 13}, {250, 130, 20, 22, 18, 21},
                          {11, 9, 9, 10, 9, 9}};
    for (const Grid &g : grids) {
        for (int rz : {1, 2, 4, 8}) {
            run_tiles_grid<float>(g, rz, 0, 0, 0, false);
            run_tiles_grid<double>(g, rz, 0, 0, 0, true);
            run_tiles_grid<float>(g, rz, 1, 4 * rz, 0, false);
        }
        for (int rz : {2, 4}) run_tiles_grid<float>(g, rz, 2, 0, 0, false);
        for (int q : {1, 2}) run_tiles_grid<float>(g, 4, 1, 16, q, false);
        run_host(g);
    }
    printf("sanitize driver: ok\n");
    return 0;
}
