// rtm_device.hip -- the plan (one simulation on HBM) and the per-step legacy
// entry points of librtm_hip.so (gfx950).
//
// A plan holds one simulation's state on HBM and runs the forward and
// backward time loops of rtmUq2D (SRC/forward_modeling.c:98-130,
// SRC/backward_modeling.c:97-126) with one fused tiled-kernel launch per
// time step, replayed from hipGraph blocks.  Everything between the
// velocity/seismogram upload and the image download stays on the device: the
// reference's disk snapshots (TEMPORARY_FILES/sourceWavefield_%05d.bin)
// become an HBM snapshot store.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <mutex>
#include <string>

#include "rtm_sim.h"

using namespace rtmd;

namespace {
thread_local std::string g_err;
}

extern "C" void rtm_set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

extern "C" const char *rtm_last_error(void) { return g_err.c_str(); }

extern "C" int rtm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------

struct rtm_plan {
    SimBuffers b;
    Exec ex;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int rz = 4;                     // tile rows per thread
    int graph_block = 100;          // steps per captured graph block (<= 1: eager)
    int interior = rtm::IN_TILES;   // interior kernel (rtm::Interior)
    int seg_rows = 0;               // rows per interior column segment
    int last_nt = 0;                // steps of the last pass
    bool last_fwd_ring = false;     // the last pass was a boundary-saving forward pass
    int last_umask = 1;             // u buffer mask of the last pass (3: pair launches)
    int fwd_src_j = -1;             // source index of the last forward pass (the reconstruction re-adds it)
    double fwd_ms = 0.0, bwd_ms = 0.0;
    int last_steps = 0;
    // shot stacks (rtm_plan_stack_image), allocated at the first stacking:
    // stack, self-image stack, and two work images for the output
    void *stk[4] = {nullptr, nullptr, nullptr, nullptr};
};

extern "C" size_t rtm_plan_footprint(const rtm_params *p, int real)
{
    if (!p || rtm_params_check(p)) return 0;
    return sim_bytes(*p, real == RTM_REAL_F64 ? 8 : 4);
}

static void plan_free(rtm_plan *q)
{
    for (void *&m : q->stk)
        if (m) { (void)hipFree(m); m = nullptr; }
    q->ex.release();
    q->b.release();
    if (q->ev0) (void)hipEventDestroy(q->ev0);
    if (q->ev1) (void)hipEventDestroy(q->ev1);
    if (q->stream) (void)hipStreamDestroy(q->stream);
}

static int plan_set_tiling(rtm_plan *q, int rz, int graph_block, int interior, int seg_rows)
{
    // the current shape stays when the new one is refused
    if (check_shape(rz, interior) || check_store_shape(q->b.p.snapshot_mode, interior) ||
        check_pair_real(q->b.real, interior))
        return -1;
    if (q->b.set_quad(interior == rtm::IN_PAIRS)) return -1;
    std::vector<int4> tiles, inner;
    append_work(tiles, inner, 0, q->b.p, rz, interior, seg_rows);
    balance_tiles_for_xcds(tiles, rtm::tile_rows(rz, interior));
    q->ex.release();
    if (q->ex.init(q->b.real, 1, (int)tiles.size(), rz, (int)inner.size(), interior)) return -1;
    q->ex.q16 = q->b.qmode == rtm::Q_FP16 || q->b.qmode == rtm::Q_BF16;
    q->ex.i16 = q->b.qmode == rtm::Q_I16;
    if (q->ex.set_tiles(tiles, q->stream) || q->ex.set_inner(inner, q->stream)) return -1;
    q->rz = rz;
    q->graph_block = graph_block;
    q->interior = interior;
    q->seg_rows = seg_rows;
    return 0;
}

extern "C" int rtm_plan_create(const rtm_params *p, int real, rtm_plan **out)
{
    if (!p || !out) { rtm_set_error("rtm_plan_create: null argument"); return -1; }
    *out = nullptr;
    if (rtm_params_check(p)) return -1;
    if (real != RTM_REAL_F32 && real != RTM_REAL_F64) {
        rtm_set_error("rtm_plan_create: unknown real type %d", real);
        return -1;
    }
    if (p->snapshot_mode < RTM_SNAP_FP32 || p->snapshot_mode > RTM_SNAP_I16) {
        rtm_set_error("rtm_plan_create: snapshot mode %d is not available", p->snapshot_mode);
        return -1;
    }
    HIP_TRY(hipSetDevice(p->device));
    rtm_plan *q = new rtm_plan();
    auto fail = [&]() { plan_free(q); delete q; return -1; };
    if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&q->ev0) != hipSuccess || hipEventCreate(&q->ev1) != hipSuccess) {
        rtm_set_error("rtm_plan_create: stream/event creation failed");
        return fail();
    }
    if (q->b.alloc(*p, real)) return fail();
    // a shape tuned earlier in this process for the same grid class, else
    // tiles of 64 x 16 ($RTM_TILE_RZ overrides the tile height)
    const char *env = getenv("RTM_TILE_RZ");
    Shape sh{env && *env ? atoi(env) : 4, rtm::IN_TILES, 0};
    if (!(env && *env)) tuner_lookup(tuner_key(real, q->b.qmode, {&q->b.p}), &sh);
    if (plan_set_tiling(q, sh.rz, 100, sh.interior, sh.seg)) return fail();
    *out = q;
    return 0;
}

extern "C" int rtm_plan_destroy(rtm_plan *q)
{
    if (!q) return 0;
    (void)hipSetDevice(q->b.p.device);
    (void)hipStreamSynchronize(q->stream);
    plan_free(q);
    delete q;
    return 0;
}

extern "C" int rtm_plan_set_tiling(rtm_plan *q, int rows_per_thread, int graph_block, int interior, int seg_rows)
{
    if (!q) { rtm_set_error("rtm_plan_set_tiling: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    HIP_TRY(hipStreamSynchronize(q->stream));
    return plan_set_tiling(q, rows_per_thread, graph_block, interior, seg_rows);
}

extern "C" int rtm_plan_set_velocity(rtm_plan *q, const float *vpad)
{
    if (!q || !vpad) { rtm_set_error("rtm_plan_set_velocity: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.set_velocity(vpad, q->stream);
}

extern "C" int rtm_plan_set_wavelet(rtm_plan *q, const float *w, int n)
{
    if (!q || !w) { rtm_set_error("rtm_plan_set_wavelet: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.set_wavelet(w, n);
}

extern "C" int rtm_plan_set_seismogram(rtm_plan *q, const float *d)
{
    if (!q || !d) { rtm_set_error("rtm_plan_set_seismogram: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.set_seismogram(d, q->stream);
}

extern "C" int rtm_plan_zero_image(rtm_plan *q)
{
    if (!q) { rtm_set_error("rtm_plan_zero_image: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    q->b.zero_image(q->stream);
    HIP_TRY(hipStreamSynchronize(q->stream));
    return 0;
}

// ---- per-point (v1) loops: injection points outside the stepped region ----

template <typename R>
static StepArgs<R> v1_args(const SimBuffers &b)
{
    const SimDesc<R> d = b.desc<R>();
    StepArgs<R> A{};
    A.X = d.X; A.Z = d.Z; A.P = d.P;
    A.TL = d.TL; A.TR = d.TR; A.TU = d.TU; A.TD = d.TD;
    A.nx = d.nx; A.nz = d.nz; A.nrec = d.nrec;
    A.h1 = d.h1; A.h2 = d.h2; A.h3 = d.h3;
    A.C = d.C;
    A.ax = d.ax; A.bx = d.bx; A.az = d.az; A.bz = d.bz;
    A.csiX = d.csiX; A.csiZ = d.csiZ;
    A.src_j = -1;
    A.rec_row = d.rec_row;
    A.recmap = d.recmap;
    A.I = d.I; A.Is = d.Is; A.Ic = d.Ic; A.Isc = d.Isc;
    return A;
}

// u at step n is u[n & 1], psi^n is psi[n & 1] (the tiled path's convention)
template <typename R>
static void v1_bind(const SimBuffers &b, StepArgs<R> &A, int n)
{
    A.u = (const R *)b.u[n & 1];
    A.U = (R *)b.u[(n + 1) & 1];
    A.psiXi = (const R *)b.psiX[n & 1];
    A.psiXo = (R *)b.psiX[(n + 1) & 1];
    A.psiZi = (const R *)b.psiZ[n & 1];
    A.psiZo = (R *)b.psiZ[(n + 1) & 1];
}

template <typename R>
static void v1_forward(rtm_plan *q, int src_j, bool src_inside, int record)
{
    const SimBuffers &b = q->b;
    const rtm_params &p = b.p;
    StepArgs<R> A = v1_args<R>(b);
    const dim3 grid((b.X + kBX - 1) / kBX, (b.Z + kBZ - 1) / kBZ), block(kBX, kBZ);
    for (int n = 0; n < p.number_of_time_steps; n++) {
        v1_bind(b, A, n);
        A.mode = rtm::M_STENCIL | rtm::M_CPML;
        const bool src = n < p.sampling_wavelet;
        if (src && src_inside) {
            A.mode |= rtm::M_SOURCE;
            A.src_j = src_j;
            A.src_w = R(b.wav_host[n]);
        }
        if (n % p.time_slices == 0) {
            A.mode |= rtm::M_SNAP;
            A.snap = (R *)b.snaps + (size_t)(n / p.time_slices) * b.nint;
        }
        hipLaunchKernelGGL(k_step<R>, grid, block, 0, q->stream, A);
        if (!src_inside) {
            // a source outside the stepped region can make the outer columns
            // non-zero, so their top-strip correction is applied too
            launch_top_edges(A, q->stream);
            if (src)
                hipLaunchKernelGGL(k_source_point<R>, dim3(1), dim3(64), 0, q->stream, A.U, A.C, src_j, b.wav, n);
        }
        if (record)
            hipLaunchKernelGGL(k_record<R>, dim3((p.number_of_receivers + 255) / 256), dim3(256), 0, q->stream,
                               (const R *)A.U, b.rec_j, p.number_of_receivers,
                               b.seis_rec + (size_t)n * p.number_of_receivers);
    }
}

template <typename R>
static void v1_backward(rtm_plan *q)
{
    const SimBuffers &b = q->b;
    const rtm_params &p = b.p;
    StepArgs<R> A = v1_args<R>(b);
    const dim3 grid((b.X + kBX - 1) / kBX, (b.Z + kBZ - 1) / kBZ), block(kBX, kBZ);
    const int TS = p.time_slices, nt = p.number_of_time_steps, nrec = p.number_of_receivers;
    for (int n = 0; n < nt; n++) {
        const int tau = nt - n - 1;
        v1_bind(b, A, n);
        A.mode = rtm::M_STENCIL | rtm::M_CPML | rtm::M_INJECT;
        A.seis_row = b.seis_t + (size_t)tau * nrec;
        // snapshot tau/TS pairs with this step; the reference's counter
        // (SRC/backward_modeling.c:96,119) equals tau/TS when nt % TS == 0
        if (tau % TS == 0) {
            A.mode |= rtm::M_IMAGE | (p.image_kahan ? rtm::M_KAHAN : 0);
            A.snap = (R *)b.snaps + (size_t)(tau / TS) * b.nint;
        }
        hipLaunchKernelGGL(k_step<R>, grid, block, 0, q->stream, A);
        if (b.n_out) {
            // receivers outside the stepped region: outer-column top-strip
            // correction, then their replacement.  None of these points is in
            // the image region, so the fused imaging above is unaffected.
            launch_top_edges(A, q->stream);
            hipLaunchKernelGGL(k_inject_list<R>, dim3((b.n_out + 255) / 256), dim3(256), 0, q->stream, A.U,
                               b.out_j, b.out_r, b.n_out, A.seis_row);
        }
    }
}

// ---- passes ----

template <typename R>
static int forward_t(rtm_plan *q, int shot_x, int shot_z, int record)
{
    SimBuffers &b = q->b;
    const rtm_params &p = b.p;
    const int xs = shot_x + p.thickness_left, zs = shot_z + p.thickness_up;
    if (xs < 0 || xs >= b.X || zs < 0 || zs >= b.Z) {
        rtm_set_error("source at padded (%d,%d) is outside the %dx%d grid", xs, zs, b.X, b.Z);
        return -1;
    }
    const bool src_inside = xs >= rtm::kHalf && xs < b.X - rtm::kHalf && zs >= rtm::kHalf && zs < b.Z - rtm::kHalf;
    const int src_j = zs * b.P + xs;
    const int nt = p.number_of_time_steps;
    b.reset_state(q->stream, b.ringmode);
    b.reset_snap_scales(q->stream);
    const bool tiled = src_inside;   // the tiled kernel covers only the stepped region
    if (!tiled && (b.qmode != rtm::Q_NONE || b.ringmode)) {
        rtm_set_error("16-bit snapshot stores and boundary saving need the source inside the stepped region");
        return -1;
    }
    q->ex.set_ring_phase(b.ringmode ? rtm::PH_FORWARD : 0, (b.NR + 255) / 256);
    HIP_TRY(hipEventRecord(q->ev0, q->stream));
    if (tiled) {
        SimDesc<R> d = b.desc<R>();
        d.phase = rtm::PH_FORWARD;
        d.flags = (b.ringmode ? rtm::F_RING : rtm::F_SNAP) | (record && !b.rec_dup && b.n_out == 0 ? rtm::F_RECORD : 0);
        d.src_j = src_j;
        if (b.ringmode) {   // the source field runs on its own buffers, kept for the backward pass
            d.ubuf[0] = (R *)b.sbuf[0];
            d.ubuf[1] = (R *)b.sbuf[1];
        }
        if (q->ex.upload_descs(&d, 1, q->stream) || q->ex.set_n0({0}, q->stream)) return -1;
        // steps that store no snapshot run the light tile instance
        q->ex.set_schedule(b.light_steps() ? rtm::PH_FORWARD : 0, nt, p.time_slices, d.flags);
        if (record && !(d.flags & rtm::F_RECORD)) {
            // receivers sharing a point or outside: record each one per step
            // (with pair items an even step's launch also steps them at n+1,
            // and the next launch completes step n+1 on the other points)
            void *const *uf = b.ringmode ? b.sbuf : b.u;   // the buffers this pass steps
            const int m = b.ringmode ? 1 : b.umask;
            for (int n = 0; n < nt; n++) {
                q->ex.launch_one(n, q->stream);
                hipLaunchKernelGGL(k_record<R>, dim3((p.number_of_receivers + 255) / 256), dim3(256), 0,
                                   q->stream, (const R *)uf[(n + 1) & m], b.rec_j, p.number_of_receivers,
                                   b.seis_rec + (size_t)n * p.number_of_receivers);
            }
        } else if (q->ex.run(nt, q->graph_block, q->stream)) {
            return -1;
        }
        if (b.flush_last_snapshot(q->stream)) return -1;
        if (q->ex.check_err(q->stream)) return -1;
    } else {
        v1_forward<R>(q, src_j, src_inside, record);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(q->ev1, q->stream));
    HIP_TRY(hipEventSynchronize(q->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, q->ev0, q->ev1));
    q->fwd_ms = ms;
    q->last_steps = nt;
    q->last_nt = nt;
    q->last_fwd_ring = b.ringmode;
    q->last_umask = tiled && !b.ringmode ? b.umask : 1;
    q->fwd_src_j = src_j;
    b.have_snaps = true;
    return 0;
}

template <typename R>
static int backward_t(rtm_plan *q)
{
    SimBuffers &b = q->b;
    const rtm_params &p = b.p;
    if (!b.have_snaps) { rtm_set_error("rtm_plan_backward: no snapshots (run forward first)"); return -1; }
    if (!b.have_seis) { rtm_set_error("rtm_plan_backward: no seismogram"); return -1; }
    const int nt = p.number_of_time_steps;
    b.reset_state(q->stream);
    q->ex.set_ring_phase(b.ringmode ? rtm::PH_BACKWARD : 0, (int)((b.nint + b.NR + 255) / 256));
    HIP_TRY(hipEventRecord(q->ev0, q->stream));
    if (b.n_out == 0) {
        SimDesc<R> d = b.desc<R>();
        d.phase = rtm::PH_BACKWARD;
        d.flags = rtm::F_IMAGE | (p.image_kahan ? rtm::F_KAHAN : 0) | (b.ringmode ? rtm::F_RING : 0);
        if (b.ringmode) d.src_j = q->fwd_src_j;   // only the reconstruction reads it in this phase
        if (q->ex.upload_descs(&d, 1, q->stream) || q->ex.set_n0({0}, q->stream)) return -1;
        // steps that image nothing run the light tile instance
        q->ex.set_schedule(b.light_steps() ? rtm::PH_BACKWARD : 0, nt, p.time_slices, d.flags);
        if (q->ex.run(nt, q->graph_block, q->stream)) return -1;
        if (q->ex.check_err(q->stream)) return -1;
    } else if (b.qmode != rtm::Q_NONE || b.ringmode) {
        rtm_set_error("16-bit snapshot stores and boundary saving need every receiver inside the stepped region");
        return -1;
    } else {
        v1_backward<R>(q);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(q->ev1, q->stream));
    HIP_TRY(hipEventSynchronize(q->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, q->ev0, q->ev1));
    q->bwd_ms = ms;
    q->last_steps = nt;
    q->last_nt = nt;
    q->last_fwd_ring = false;
    q->last_umask = b.n_out == 0 ? b.umask : 1;
    if (b.ringmode) b.have_snaps = false;   // the reconstruction consumed the saved source field
    return 0;
}

static int ready(rtm_plan *q, const char *who)
{
    if (!q) { rtm_set_error("%s: null plan", who); return -1; }
    if (!q->b.have_velocity) { rtm_set_error("%s: velocity not set", who); return -1; }
    if (hipSetDevice(q->b.p.device) != hipSuccess) { rtm_set_error("%s: hipSetDevice failed", who); return -1; }
    return 0;
}

extern "C" int rtm_plan_forward(rtm_plan *q, int shot_x, int shot_z, int record)
{
    if (ready(q, "rtm_plan_forward")) return -1;
    if (q->b.wav_host.size() != (size_t)q->b.p.sampling_wavelet) {
        rtm_set_error("rtm_plan_forward: wavelet not set");
        return -1;
    }
    return q->b.real == RTM_REAL_F64 ? forward_t<double>(q, shot_x, shot_z, record)
                                     : forward_t<float>(q, shot_x, shot_z, record);
}

extern "C" int rtm_plan_backward(rtm_plan *q)
{
    if (ready(q, "rtm_plan_backward")) return -1;
    return q->b.real == RTM_REAL_F64 ? backward_t<double>(q) : backward_t<float>(q);
}

extern "C" int rtm_plan_migrate(rtm_plan *q, int shot_x, int shot_z)
{
    int rc = rtm_plan_forward(q, shot_x, shot_z, 0);
    if (rc == 0) rc = rtm_plan_backward(q);
    return rc;
}

extern "C" int rtm_plan_get_seismogram(rtm_plan *q, float *d)
{
    if (!q || !d) { rtm_set_error("rtm_plan_get_seismogram: null argument"); return -1; }
    const int nrec = q->b.p.number_of_receivers, nt = q->b.p.number_of_time_steps;
    std::vector<float> t((size_t)nt * nrec);
    HIP_TRY(hipSetDevice(q->b.p.device));
    HIP_TRY(hipStreamSynchronize(q->stream));
    HIP_TRY(hipMemcpy(t.data(), q->b.seis_rec, sizeof(float) * t.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < nrec; r++)
        for (int n = 0; n < nt; n++) d[(size_t)r * nt + n] = t[(size_t)n * nrec + r];
    return 0;
}

template <typename R, typename O>
static int get_image_t(rtm_plan *q, O *I, O *Is)
{
    const SimBuffers &b = q->b;
    const int nx = b.p.nx, nz = b.p.nz;
    std::vector<R> h(b.nint);
    HIP_TRY(hipStreamSynchronize(q->stream));
    void *src[2] = {b.I, b.Is};
    O *dst[2] = {I, Is};
    for (int t = 0; t < 2; t++) {
        if (!dst[t]) continue;
        HIP_TRY(hipMemcpy(h.data(), src[t], sizeof(R) * b.nint, hipMemcpyDeviceToHost));
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) dst[t][(size_t)i * nz + k] = O(h[(size_t)k * nx + i]);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// shot stacking on the device (SRC/rtm_iso2d.c:115-152)
// ---------------------------------------------------------------------------

namespace {

// stack += I and stack_self += Is, one shot's images, in FP of the plan
template <typename R>
__global__ void k_stack_add(R *__restrict__ st, R *__restrict__ ss, const R *__restrict__ I,
                            const R *__restrict__ Is, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        st[q] = st[q] + I[q];
        ss[q] = ss[q] + Is[q];
    }
}

// the stack, or with normalize its illumination-normalized form stack /
// stack_self (0 where stack_self is 0; the division the reference leaves
// commented out, SRC/rtm_iso2d.c:150)
template <typename R>
__global__ void k_stack_out(R *__restrict__ out, const R *__restrict__ st, const R *__restrict__ ss, size_t n,
                            int normalize)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
        out[q] = normalize ? (ss[q] != R(0) ? st[q] / ss[q] : R(0)) : st[q];
}

// 8th-order Laplacian without coefficient or spacing over [4,Nx-4) x
// [4,Nz-4), zero elsewhere (SRC/math_functions.c:15-24), on the device's
// [Nz][Nx] image layout: the x neighbours are q -+ m, the z neighbours
// q -+ m*Nx.  Each ring sums (x-, x+, z-, z+) in that order and the rings
// are added from m = 0 up, as the reference writes it.
template <typename R>
__global__ void k_laplacian(R *__restrict__ lp, const R *__restrict__ c, int nx, int nz)
{
    const size_t n = (size_t)nx * nz, stride = (size_t)gridDim.x * blockDim.x;
    const R w0 = R(2.0f * -2.97399944f), w1 = R(1.70507669f), w2 = R(-0.25861812f), w3 = R(0.04577745f),
            w4 = R(-0.00523630f);
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const int i = (int)(q % (size_t)nx), k = (int)(q / (size_t)nx);
        if (i < 4 || i >= nx - 4 || k < 4 || k >= nz - 4) {
            lp[q] = R(0);
            continue;
        }
        const size_t X1 = 1, Z1 = (size_t)nx;
        const R t1 = ((c[q - X1] + c[q + X1]) + c[q - Z1]) + c[q + Z1];
        const R t2 = ((c[q - 2 * X1] + c[q + 2 * X1]) + c[q - 2 * Z1]) + c[q + 2 * Z1];
        const R t3 = ((c[q - 3 * X1] + c[q + 3 * X1]) + c[q - 3 * Z1]) + c[q + 3 * Z1];
        const R t4 = ((c[q - 4 * X1] + c[q + 4 * X1]) + c[q - 4 * Z1]) + c[q + 4 * Z1];
        lp[q] = (((w0 * c[q] + w1 * t1) + w2 * t2) + w3 * t3) + w4 * t4;
    }
}

int stack_alloc(rtm_plan *q)
{
    if (q->stk[0]) return 0;
    const size_t bytes = q->b.nint * q->b.esz;
    for (void *&m : q->stk) {
        if (hipMalloc(&m, bytes) != hipSuccess) {
            rtm_set_error("rtm_plan_stack_image: hipMalloc(%zu) failed", bytes);
            return -1;
        }
    }
    return 0;
}

template <typename R>
int zero_stack_t(rtm_plan *q)
{
    for (int t = 0; t < 2; t++) {
        hipLaunchKernelGGL(k_fill_zero<R>, dim3(1024), dim3(256), 0, q->stream, (R *)q->stk[t], q->b.nint);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

template <typename R>
int stack_image_t(rtm_plan *q)
{
    const bool fresh = !q->stk[0];
    if (stack_alloc(q)) return -1;
    if (fresh && zero_stack_t<R>(q)) return -1;
    hipLaunchKernelGGL(k_stack_add<R>, dim3(1024), dim3(256), 0, q->stream, (R *)q->stk[0], (R *)q->stk[1],
                       (const R *)q->b.I, (const R *)q->b.Is, q->b.nint);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename R>
int get_stack_t(rtm_plan *q, int normalize, float *stack, float *lap)
{
    const int nx = q->b.p.nx, nz = q->b.p.nz;
    if (stack_alloc(q)) return -1;
    R *out = (R *)q->stk[2], *lp = (R *)q->stk[3];
    hipLaunchKernelGGL(k_stack_out<R>, dim3(1024), dim3(256), 0, q->stream, out, (const R *)q->stk[0],
                       (const R *)q->stk[1], q->b.nint, normalize);
    hipLaunchKernelGGL(k_laplacian<R>, dim3(1024), dim3(256), 0, q->stream, lp, (const R *)out, nx, nz);
    HIP_TRY(hipGetLastError());
    std::vector<R> h(q->b.nint);
    void *src[2] = {out, lp};
    float *dst[2] = {stack, lap};
    HIP_TRY(hipStreamSynchronize(q->stream));
    for (int t = 0; t < 2; t++) {
        if (!dst[t]) continue;
        HIP_TRY(hipMemcpy(h.data(), src[t], sizeof(R) * q->b.nint, hipMemcpyDeviceToHost));
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) dst[t][(size_t)i * nz + k] = float(h[(size_t)k * nx + i]);
    }
    return 0;
}

}  // namespace

extern "C" int rtm_plan_stack_image(rtm_plan *q)
{
    if (!q) { rtm_set_error("rtm_plan_stack_image: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.real == RTM_REAL_F64 ? stack_image_t<double>(q) : stack_image_t<float>(q);
}

extern "C" int rtm_plan_zero_stack(rtm_plan *q)
{
    if (!q) { rtm_set_error("rtm_plan_zero_stack: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    if (stack_alloc(q)) return -1;
    return q->b.real == RTM_REAL_F64 ? zero_stack_t<double>(q) : zero_stack_t<float>(q);
}

extern "C" int rtm_plan_get_stack(rtm_plan *q, int normalize, float *stack, float *laplacian)
{
    if (!q) { rtm_set_error("rtm_plan_get_stack: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    if (!q->stk[0] && rtm_plan_zero_stack(q)) return -1;   // nothing stacked: zero images
    return q->b.real == RTM_REAL_F64 ? get_stack_t<double>(q, normalize, stack, laplacian)
                                     : get_stack_t<float>(q, normalize, stack, laplacian);
}

extern "C" int rtm_plan_get_image(rtm_plan *q, float *I, float *Is)
{
    if (!q || !I) { rtm_set_error("rtm_plan_get_image: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.real == RTM_REAL_F64 ? get_image_t<double, float>(q, I, Is) : get_image_t<float, float>(q, I, Is);
}

extern "C" int rtm_plan_get_image_f64(rtm_plan *q, double *I, double *Is)
{
    if (!q || !I) { rtm_set_error("rtm_plan_get_image_f64: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.real == RTM_REAL_F64 ? get_image_t<double, double>(q, I, Is) : get_image_t<float, double>(q, I, Is);
}

template <typename R, typename O>
static int get_fields_t(rtm_plan *q, O *u_prev, O *u_last)
{
    const SimBuffers &b = q->b;
    std::vector<R> h(b.nfield);
    HIP_TRY(hipStreamSynchronize(q->stream));
    // after nt steps u[nt & m] holds u^{nt} and u[(nt-1) & m] u^{nt-1}
    const int nt = q->last_nt, m = q->last_umask;
    void *const *w = q->last_fwd_ring ? b.sbuf : b.u;
    void *src[2] = {w[(nt - 1) & m], w[nt & m]};
    O *dst[2] = {u_prev, u_last};
    for (int t = 0; t < 2; t++) {
        if (!dst[t]) continue;
        HIP_TRY(hipMemcpy(h.data(), src[t], sizeof(R) * b.nfield, hipMemcpyDeviceToHost));
        from_device_layout(h.data(), dst[t], b.X, b.Z, b.P);
    }
    return 0;
}

extern "C" int rtm_plan_get_fields(rtm_plan *q, float *u_prev, float *u_last)
{
    if (!q) { rtm_set_error("rtm_plan_get_fields: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.real == RTM_REAL_F64 ? get_fields_t<double, float>(q, u_prev, u_last)
                                     : get_fields_t<float, float>(q, u_prev, u_last);
}

extern "C" int rtm_plan_get_fields_f64(rtm_plan *q, double *u_prev, double *u_last)
{
    if (!q) { rtm_set_error("rtm_plan_get_fields_f64: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    return q->b.real == RTM_REAL_F64 ? get_fields_t<double, double>(q, u_prev, u_last)
                                     : get_fields_t<float, double>(q, u_prev, u_last);
}

extern "C" int rtm_plan_synchronize(rtm_plan *q)
{
    if (!q) { rtm_set_error("rtm_plan_synchronize: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    HIP_TRY(hipStreamSynchronize(q->stream));
    return 0;
}

// Autotuner: every candidate shape timed on this plan's grid; the fastest is
// kept and cached for the grid class.
template <typename R>
static int plan_autotune_t(rtm_plan *q, int steps, int cap, int *shapes, double *ms, int *n_out)
{
    const auto &cand = tuner_candidates();
    double best_t = 1e30;
    Shape best = cand[0];
    int k = 0;
    for (const Shape &c : cand) {
        // not a shape of this store or element type
        if (check_store_shape(q->b.p.snapshot_mode, c.interior) || check_pair_real(q->b.real, c.interior)) continue;
        if (plan_set_tiling(q, c.rz, q->graph_block, c.interior, c.seg)) return -1;
        double t = 0.0;
        if (time_steps<R>(q->ex, {q->b.desc<R>()}, steps, q->graph_block, q->stream, &t)) return -1;
        if (k < cap) {
            if (shapes) {
                shapes[4 * k] = c.rz; shapes[4 * k + 1] = q->graph_block;
                shapes[4 * k + 2] = c.interior; shapes[4 * k + 3] = c.seg;
            }
            if (ms) ms[k] = t;
        }
        k++;
        if (t < best_t) { best_t = t; best = c; }
    }
    if (n_out) *n_out = k < cap ? k : cap;
    if (plan_set_tiling(q, best.rz, q->graph_block, best.interior, best.seg)) return -1;
    tuner_store(tuner_key(q->b.real, q->b.qmode, {&q->b.p}), best);
    return 0;
}

extern "C" int rtm_plan_autotune(rtm_plan *q, int steps, int cap, int *shapes, double *ms_per_step, int *n_out)
{
    if (!q || steps < 1) { rtm_set_error("rtm_plan_autotune: bad argument"); return -1; }
    HIP_TRY(hipSetDevice(q->b.p.device));
    HIP_TRY(hipStreamSynchronize(q->stream));
    return q->b.real == RTM_REAL_F64 ? plan_autotune_t<double>(q, steps, cap, shapes, ms_per_step, n_out)
                                     : plan_autotune_t<float>(q, steps, cap, shapes, ms_per_step, n_out);
}

extern "C" int rtm_plan_get_tiling(rtm_plan *q, int *rows_per_thread, int *graph_block, int *interior, int *seg_rows)
{
    if (!q) { rtm_set_error("rtm_plan_get_tiling: null plan"); return -1; }
    if (rows_per_thread) *rows_per_thread = q->rz;
    if (graph_block) *graph_block = q->graph_block;
    if (interior) *interior = q->interior;
    if (seg_rows) *seg_rows = q->seg_rows;
    return 0;
}

extern "C" int rtm_plan_last_timing(rtm_plan *q, double *fwd_step_ms, double *bwd_step_ms, int *steps)
{
    if (!q) { rtm_set_error("rtm_plan_last_timing: null plan"); return -1; }
    const double n = q->last_steps ? (double)q->last_steps : 1.0;
    if (fwd_step_ms) *fwd_step_ms = q->fwd_ms / n;
    if (bwd_step_ms) *bwd_step_ms = q->bwd_ms / n;
    if (steps) *steps = q->last_steps;
    return 0;
}

// ---------------------------------------------------------------------------
// snapshot-store table keyed by simulation id
// ---------------------------------------------------------------------------

static std::mutex g_table_mu;
static std::map<int, rtm_plan *> g_table;

extern "C" int rtmd_table_put(int sim_id, rtm_plan *plan)
{
    rtm_plan *old = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        auto it = g_table.find(sim_id);
        if (it != g_table.end()) old = it->second;
        g_table[sim_id] = plan;
    }
    if (old && old != plan) rtm_plan_destroy(old);
    return 0;
}

extern "C" rtm_plan *rtmd_table_take(int sim_id)
{
    std::lock_guard<std::mutex> lk(g_table_mu);
    auto it = g_table.find(sim_id);
    if (it == g_table.end()) return nullptr;
    rtm_plan *p = it->second;
    g_table.erase(it);
    return p;
}

// ---------------------------------------------------------------------------
// legacy per-step symbols: host pointers in the reference layout
// ---------------------------------------------------------------------------

namespace {

struct DevBufs {
    std::vector<void *> ptrs;
    ~DevBufs() { for (void *p : ptrs) (void)hipFree(p); }
    template <typename T>
    T *alloc(size_t n)
    {
        void *p = nullptr;
        if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) return nullptr;
        (void)hipMemset(p, 0, (n ? n : 1) * sizeof(T));
        ptrs.push_back(p);
        return (T *)p;
    }
};

struct Geo {
    int X, Z, P, TL, TR, TU, TD;
};

Geo geo_of(const rtm_params *p)
{
    Geo g;
    g.TL = p->thickness_left; g.TR = p->thickness_right; g.TU = p->thickness_up; g.TD = p->thickness_down;
    g.X = g.TL + p->nx + g.TR;
    g.Z = g.TU + p->nz + g.TD;
    g.P = round_up(g.X, 64);
    return g;
}

int put_field(float *dev, const float *ref, const Geo &g)
{
    std::vector<float> h((size_t)g.P * g.Z);
    to_device_layout(ref, h.data(), g.X, g.Z, g.P);
    HIP_TRY(hipMemcpy(dev, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

int get_field(float *ref, const float *dev, const Geo &g)
{
    std::vector<float> h((size_t)g.P * g.Z);
    HIP_TRY(hipMemcpy(h.data(), dev, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    from_device_layout(h.data(), ref, g.X, g.Z, g.P);
    return 0;
}

StepArgs<float> legacy_args(const rtm_params *p, const Geo &g)
{
    StepArgs<float> A{};
    A.X = g.X; A.Z = g.Z; A.P = g.P;
    A.TL = g.TL; A.TR = g.TR; A.TU = g.TU; A.TD = g.TD;
    A.nx = p->nx; A.nz = p->nz; A.nrec = p->number_of_receivers;
    const float h = p->spacing_grid;
    A.h1 = 1.0f / h; A.h2 = h * h; A.h3 = 1.0f / (h * h);
    A.src_j = -1;
    A.rec_row = -1;
    return A;
}

}  // namespace

extern "C" int rtmd_stencil_host(const rtm_params *p, const float *pond, const float *u, float *U)
{
    if (rtm_params_check(p)) return -1;
    HIP_TRY(hipSetDevice(p->device));
    const Geo g = geo_of(p);
    const size_t nf = (size_t)g.P * g.Z;
    DevBufs b;
    float *dC = b.alloc<float>(nf), *du = b.alloc<float>(nf), *dU = b.alloc<float>(nf);
    if (!dC || !du || !dU) { rtm_set_error("iso_acoustic_wave_equation: hipMalloc failed"); return -1; }
    if (put_field(dC, pond, g) || put_field(du, u, g) || put_field(dU, U, g)) return -1;
    StepArgs<float> A = legacy_args(p, g);
    A.mode = rtm::M_STENCIL;
    A.u = du; A.U = dU; A.C = dC;
    hipLaunchKernelGGL(k_step<float>, dim3((g.X + kBX - 1) / kBX, (g.Z + kBZ - 1) / kBZ), dim3(kBX, kBZ), 0, 0, A);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return get_field(U, dU, g);
}

extern "C" int rtmd_cpml_host(const rtm_params *p, const float *u, float *U, const float *C,
                              const float *aL, const float *bL, const float *aR, const float *bR,
                              const float *aT, const float *bT, const float *aB, const float *bB,
                              float *psiL, float *psiR, float *psiT, float *psiB,
                              float *csiL, float *csiR, float *csiT, float *csiB)
{
    if (rtm_params_check(p)) return -1;
    HIP_TRY(hipSetDevice(p->device));
    const Geo g = geo_of(p);
    const size_t nf = (size_t)g.P * g.Z;
    const int XR = g.X - g.TR, ZD = g.Z - g.TD;
    DevBufs b;
    float *dC = b.alloc<float>(nf), *du = b.alloc<float>(nf), *dU = b.alloc<float>(nf);
    float *pXi = b.alloc<float>(nf), *pXo = b.alloc<float>(nf), *pZi = b.alloc<float>(nf),
          *pZo = b.alloc<float>(nf), *cX = b.alloc<float>(nf), *cZ = b.alloc<float>(nf);
    float *ax = b.alloc<float>(g.X), *bx = b.alloc<float>(g.X), *az = b.alloc<float>(g.Z),
          *bz = b.alloc<float>(g.Z);
    if (!dC || !du || !dU || !pXi || !pXo || !pZi || !pZo || !cX || !cZ || !ax || !bx || !az || !bz) {
        rtm_set_error("aiso_attenuation_cpml: hipMalloc failed");
        return -1;
    }
    if (put_field(dC, C, g) || put_field(du, u, g) || put_field(dU, U, g)) return -1;
    // strips -> full-field device layout (strip layouts of
    // SRC/absorbingBoundaryCondition.c:180,226,272,319)
    std::vector<float> hpX(nf, 0.f), hpZ(nf, 0.f), hcX(nf, 0.f), hcZ(nf, 0.f);
    for (int i = 0; i < g.TL; i++)
        for (int k = 0; k < g.Z; k++) {
            hpX[(size_t)k * g.P + i] = psiL[(size_t)g.Z * i + k];
            hcX[(size_t)k * g.P + i] = csiL[(size_t)g.Z * i + k];
        }
    for (int i = XR; i < g.X; i++)
        for (int k = 0; k < g.Z; k++) {
            hpX[(size_t)k * g.P + i] = psiR[(size_t)g.Z * (i - XR) + k];
            hcX[(size_t)k * g.P + i] = csiR[(size_t)g.Z * (i - XR) + k];
        }
    for (int i = 0; i < g.X; i++) {
        for (int k = 0; k < g.TU; k++) {
            hpZ[(size_t)k * g.P + i] = psiT[(size_t)g.TU * i + k];
            hcZ[(size_t)k * g.P + i] = csiT[(size_t)g.TU * i + k];
        }
        for (int k = ZD; k < g.Z; k++) {
            hpZ[(size_t)k * g.P + i] = psiB[(size_t)g.TD * i + (k - ZD)];
            hcZ[(size_t)k * g.P + i] = csiB[(size_t)g.TD * i + (k - ZD)];
        }
    }
    std::vector<float> hax(g.X, 0.f), hbx(g.X, 0.f), haz(g.Z, 0.f), hbz(g.Z, 0.f);
    for (int i = 0; i < g.TL; i++) { hax[i] = aL[i]; hbx[i] = bL[i]; }
    for (int i = 0; i < g.TR; i++) { hax[XR + i] = aR[i]; hbx[XR + i] = bR[i]; }
    for (int k = 0; k < g.TU; k++) { haz[k] = aT[k]; hbz[k] = bT[k]; }
    for (int k = 0; k < g.TD; k++) { haz[ZD + k] = aB[k]; hbz[ZD + k] = bB[k]; }
    HIP_TRY(hipMemcpy(pXi, hpX.data(), sizeof(float) * nf, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pXo, hpX.data(), sizeof(float) * nf, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pZi, hpZ.data(), sizeof(float) * nf, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pZo, hpZ.data(), sizeof(float) * nf, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(cX, hcX.data(), sizeof(float) * nf, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(cZ, hcZ.data(), sizeof(float) * nf, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ax, hax.data(), sizeof(float) * g.X, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bx, hbx.data(), sizeof(float) * g.X, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(az, haz.data(), sizeof(float) * g.Z, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(bz, hbz.data(), sizeof(float) * g.Z, hipMemcpyHostToDevice));
    StepArgs<float> A = legacy_args(p, g);
    A.mode = rtm::M_CPML;
    A.u = du; A.U = dU; A.C = dC;
    A.ax = ax; A.bx = bx; A.az = az; A.bz = bz;
    A.psiXi = pXi; A.psiXo = pXo; A.csiX = cX;
    A.psiZi = pZi; A.psiZo = pZo; A.csiZ = cZ;
    hipLaunchKernelGGL(k_step<float>, dim3((g.X + kBX - 1) / kBX, (g.Z + kBZ - 1) / kBZ), dim3(kBX, kBZ), 0, 0, A);
    launch_top_edges(A, 0);   // full-width top pass, SRC/absorbingBoundaryCondition.c:239
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    if (get_field(U, dU, g)) return -1;
    HIP_TRY(hipMemcpy(hpX.data(), pXo, sizeof(float) * nf, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hpZ.data(), pZo, sizeof(float) * nf, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hcX.data(), cX, sizeof(float) * nf, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hcZ.data(), cZ, sizeof(float) * nf, hipMemcpyDeviceToHost));
    for (int i = 0; i < g.TL; i++)
        for (int k = 0; k < g.Z; k++) {
            psiL[(size_t)g.Z * i + k] = hpX[(size_t)k * g.P + i];
            csiL[(size_t)g.Z * i + k] = hcX[(size_t)k * g.P + i];
        }
    for (int i = XR; i < g.X; i++)
        for (int k = 0; k < g.Z; k++) {
            psiR[(size_t)g.Z * (i - XR) + k] = hpX[(size_t)k * g.P + i];
            csiR[(size_t)g.Z * (i - XR) + k] = hcX[(size_t)k * g.P + i];
        }
    for (int i = 0; i < g.X; i++) {
        for (int k = 0; k < g.TU; k++) {
            psiT[(size_t)g.TU * i + k] = hpZ[(size_t)k * g.P + i];
            csiT[(size_t)g.TU * i + k] = hcZ[(size_t)k * g.P + i];
        }
        for (int k = ZD; k < g.Z; k++) {
            psiB[(size_t)g.TD * i + (k - ZD)] = hpZ[(size_t)k * g.P + i];
            csiB[(size_t)g.TD * i + (k - ZD)] = hcZ[(size_t)k * g.P + i];
        }
    }
    return 0;
}

extern "C" int rtmd_image_host(const rtm_params *p, float *I, float *Is, const float *S, const float *R)
{
    if (rtm_params_check(p)) return -1;
    HIP_TRY(hipSetDevice(p->device));
    const Geo g = geo_of(p);
    const size_t nf = (size_t)g.P * g.Z, ni = (size_t)p->nx * p->nz;
    DevBufs b;
    float *dS = b.alloc<float>(nf), *dR = b.alloc<float>(nf), *dI = b.alloc<float>(ni),
          *dIs = Is ? b.alloc<float>(ni) : nullptr;
    if (!dS || !dR || !dI || (Is && !dIs)) { rtm_set_error("imaging: hipMalloc failed"); return -1; }
    if (put_field(dS, S, g) || put_field(dR, R, g)) return -1;
    std::vector<float> h(ni);
    auto up = [&](float *dev, const float *ref) -> int {
        for (int i = 0; i < p->nx; i++)
            for (int k = 0; k < p->nz; k++) h[(size_t)k * p->nx + i] = ref[(size_t)i * p->nz + k];
        HIP_TRY(hipMemcpy(dev, h.data(), sizeof(float) * ni, hipMemcpyHostToDevice));
        return 0;
    };
    auto down = [&](float *ref, const float *dev) -> int {
        HIP_TRY(hipMemcpy(h.data(), dev, sizeof(float) * ni, hipMemcpyDeviceToHost));
        for (int i = 0; i < p->nx; i++)
            for (int k = 0; k < p->nz; k++) ref[(size_t)i * p->nz + k] = h[(size_t)k * p->nx + i];
        return 0;
    };
    if (up(dI, I) || (Is && up(dIs, Is))) return -1;
    hipLaunchKernelGGL(k_image_padded<float>, dim3((p->nx + 255) / 256, p->nz), dim3(256), 0, 0, dI, dIs,
                       dS, dR, p->nx, p->nz, g.TL, g.TU, g.P);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    if (down(I, dI) || (Is && down(Is, dIs))) return -1;
    return 0;
}
