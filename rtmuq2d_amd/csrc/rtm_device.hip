// rtm_device.hip -- HIP device layer of librtm_hip.so (gfx950).
//
// Holds one simulation's state on HBM (rtm_plan) and runs the forward and
// backward time loops of rtmUq2D (SRC/forward_modeling.c:98-130,
// SRC/backward_modeling.c:97-126) with one fused kernel launch per time step.
// Everything between the velocity/seismogram upload and the image download
// stays on the device: the reference's disk snapshots
// (TEMPORARY_FILES/sourceWavefield_%05d.bin) become an HBM snapshot store.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "rtm_internal.h"
#include "rtm_step.h"

using rtm::StepArgs;

namespace {

thread_local std::string g_err;

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            rtm_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                          __FILE__, __LINE__);                                         \
            return -1;                                                                 \
        }                                                                              \
    } while (0)

constexpr int kBX = 64;   // one wavefront along x
constexpr int kBZ = 4;

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

template <typename R>
__global__ void __launch_bounds__(kBX * kBZ) k_step(StepArgs<R> A)
{
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int k = blockIdx.y * kBZ + threadIdx.y;
    if (i < rtm::kHalf || i >= A.X - rtm::kHalf || k < rtm::kHalf || k >= A.Z - rtm::kHalf) return;
    rtm::step_point(A, i, k);
}

// Top-strip correction of the outer columns i < 4 and i >= X-4
// (rtm::cpml_top_edge_point); 8 columns x (TU-8) rows.
template <typename R>
__global__ void k_top_edges(StepArgs<R> A)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = A.TU - 2 * rtm::kHalf;
    if (t >= 8 * rows) return;
    const int c = t % 8, k = rtm::kHalf + t / 8;
    const int i = c < 4 ? c : A.X - 8 + c;
    rtm::cpml_top_edge_point(A, i, k);
}

template <typename R>
static void launch_top_edges(const StepArgs<R> &A, hipStream_t s)
{
    const int n = 8 * (A.TU - 2 * rtm::kHalf);
    hipLaunchKernelGGL(k_top_edges<R>, dim3((n + 63) / 64), dim3(64), 0, s, A);
}

// coef = (v*aux)*(v*aux) on the padded velocity (already in device layout),
// SRC/forward_modeling.c:79.
template <typename R>
__global__ void k_coef(const float *__restrict__ v, R *__restrict__ C, R aux, int X, int Z, int P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= X || k >= Z) return;
    const size_t j = (size_t)k * P + i;
    const R vv = R(v[j]);
    C[j] = (vv * aux) * (vv * aux);
}

// Points the fused kernel does not cover (outside 4 <= i < X-4, 4 <= k < Z-4):
// receiver replacement, SRC/backward_modeling.c:106-110.
template <typename R>
__global__ void k_inject_list(R *__restrict__ U, const int *__restrict__ jlist,
                              const int *__restrict__ rlist, int n, const float *__restrict__ seis_row)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) U[jlist[t]] = R(seis_row[rlist[t]]);
}

// ... and source add, SRC/forward_modeling.c:105-109.
template <typename R>
__global__ void k_source_point(R *__restrict__ U, const R *__restrict__ C, int j, R w)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) U[j] = U[j] + w * C[j];
}

// Seismogram recording (forward): d[n*nrec + r] = U at receiver r.
template <typename R>
__global__ void k_record(const R *__restrict__ U, const int *__restrict__ rec_j, int nrec,
                         float *__restrict__ row)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) row[r] = float(U[rec_j[r]]);
}

// Imaging on padded fields (legacy normalized_cross_correlation),
// SRC/image_conditions.c:20-33; image layout [nz][nx] x fastest.
template <typename R>
__global__ void k_image_padded(R *__restrict__ I, R *__restrict__ Is, const R *__restrict__ S,
                               const R *__restrict__ Rw, int nx, int nz, int TL, int TU, int P)
{
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    const int kk = blockIdx.y;
    if (ii >= nx || kk >= nz) return;
    const size_t j = (size_t)(kk + TU) * P + (ii + TL);
    const size_t q = (size_t)kk * nx + ii;
    I[q] = I[q] + S[j] * Rw[j];
    if (Is) Is[q] = Is[q] + S[j] * S[j];
}

// ---------------------------------------------------------------------------
// host-side layout helpers (reference [X][Z] z-fastest <-> device [Z][P])
// ---------------------------------------------------------------------------

template <typename Src, typename Dst>
void to_device_layout(const Src *ref, Dst *dev, int X, int Z, int P)
{
    for (int k = 0; k < Z; k++) {
        Dst *row = dev + (size_t)k * P;
        for (int i = 0; i < X; i++) row[i] = Dst(ref[(size_t)i * Z + k]);
        for (int i = X; i < P; i++) row[i] = Dst(0);
    }
}

template <typename Src, typename Dst>
void from_device_layout(const Src *dev, Dst *ref, int X, int Z, int P)
{
    for (int i = 0; i < X; i++)
        for (int k = 0; k < Z; k++) ref[(size_t)i * Z + k] = Dst(dev[(size_t)k * P + i]);
}

}  // namespace

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------

struct rtm_plan {
    rtm_params p{};
    int real = RTM_REAL_F32;
    size_t esz = 4;
    int X = 0, Z = 0, P = 0;
    size_t nfield = 0, nint = 0, nsnap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *u[2] = {nullptr, nullptr};
    void *C = nullptr;
    void *psiX[2] = {nullptr, nullptr}, *psiZ[2] = {nullptr, nullptr};
    void *csiX = nullptr, *csiZ = nullptr;
    void *ax = nullptr, *bx = nullptr, *az = nullptr, *bz = nullptr;
    int *recmap = nullptr;          // [X]
    int *rec_j = nullptr;           // [nrec] device index of each receiver
    int *out_j = nullptr, *out_r = nullptr;   // receivers outside the fused region
    int n_out = 0;
    float *wav = nullptr;           // [n_w]
    float *seis_t = nullptr;        // [nt][nrec] (backward input, t-major)
    float *seis_rec = nullptr;      // [nt][nrec] (forward recording)
    void *snaps = nullptr;          // [nsnap][nz][nx]
    void *I = nullptr, *Is = nullptr, *Ic = nullptr, *Isc = nullptr;
    std::vector<float> wav_host;
    int cur = 0;                    // u[cur] = u^n after a loop
    int psi_cur = 0;
    double fwd_ms = 0.0, bwd_ms = 0.0;   // per-pass event time
    int last_steps = 0;
    bool have_velocity = false, have_seis = false, have_snaps = false;
};

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

static size_t plan_bytes(const rtm_params &p, size_t esz, size_t *nsnap_out)
{
    const int X = p.thickness_left + p.nx + p.thickness_right;
    const int Z = p.thickness_up + p.nz + p.thickness_down;
    const size_t P = (size_t)round_up(X, 64);
    const size_t nfield = P * Z;
    const size_t nint = (size_t)p.nx * p.nz;
    const size_t nsnap = (size_t)(p.number_of_time_steps / p.time_slices);
    if (nsnap_out) *nsnap_out = nsnap;
    size_t b = 0;
    b += 9 * nfield * esz;                          // u0,u1,C,psiX x2,psiZ x2,csiX,csiZ
    b += (size_t)(2 * X + 2 * Z) * esz;             // profiles
    b += (size_t)X * 4 + 3 * (size_t)p.number_of_receivers * 4;
    b += (size_t)p.sampling_wavelet * 4;
    b += 2 * (size_t)p.number_of_time_steps * p.number_of_receivers * 4;
    b += nsnap * nint * esz;
    b += 4 * nint * esz;                            // I, Is, Ic, Isc
    return b;
}

extern "C" size_t rtm_plan_footprint(const rtm_params *p, int real)
{
    if (!p || rtm_params_check(p)) return 0;
    return plan_bytes(*p, real == RTM_REAL_F64 ? 8 : 4, nullptr);
}

static void plan_free(rtm_plan *q)
{
    void *all[] = {q->u[0], q->u[1], q->C, q->psiX[0], q->psiX[1], q->psiZ[0], q->psiZ[1],
                   q->csiX, q->csiZ, q->ax, q->bx, q->az, q->bz, q->recmap, q->rec_j, q->out_j,
                   q->out_r, q->wav, q->seis_t, q->seis_rec, q->snaps, q->I, q->Is, q->Ic, q->Isc};
    for (void *ptr : all)
        if (ptr) (void)hipFree(ptr);
    if (q->ev0) (void)hipEventDestroy(q->ev0);
    if (q->ev1) (void)hipEventDestroy(q->ev1);
    if (q->stream) (void)hipStreamDestroy(q->stream);
}

template <typename R>
static int upload_profiles(rtm_plan *q)
{
    const rtm_params &p = q->p;
    const int TL = p.thickness_left, TR = p.thickness_right, TU = p.thickness_up,
              TD = p.thickness_down;
    std::vector<R> aL(TL), aR(TR), aT(TU), aB(TD), bL(TL), bR(TR), bT(TU), bB(TD);
    if constexpr (sizeof(R) == 4)
        rtm_profiles_f32(&p, aL.data(), aR.data(), aT.data(), aB.data(), bL.data(), bR.data(),
                         bT.data(), bB.data());
    else
        rtm_profiles_f64(&p, aL.data(), aR.data(), aT.data(), aB.data(), bL.data(), bR.data(),
                         bT.data(), bB.data());
    // full-length profiles: x strips use the left/right arrays, z strips the
    // top/bottom arrays, indexed as SRC/absorbingBoundaryCondition.c:181,273,227,320
    std::vector<R> ax(q->X, R(0)), bx(q->X, R(0)), az(q->Z, R(0)), bz(q->Z, R(0));
    for (int i = 0; i < TL; i++) { ax[i] = aL[i]; bx[i] = bL[i]; }
    for (int i = 0; i < TR; i++) { ax[q->X - TR + i] = aR[i]; bx[q->X - TR + i] = bR[i]; }
    for (int k = 0; k < TU; k++) { az[k] = aT[k]; bz[k] = bT[k]; }
    for (int k = 0; k < TD; k++) { az[q->Z - TD + k] = aB[k]; bz[q->Z - TD + k] = bB[k]; }
    HIP_TRY(hipMemcpy(q->ax, ax.data(), sizeof(R) * q->X, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(q->bx, bx.data(), sizeof(R) * q->X, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(q->az, az.data(), sizeof(R) * q->Z, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(q->bz, bz.data(), sizeof(R) * q->Z, hipMemcpyHostToDevice));
    return 0;
}

// receiver tables: fused map on the receiver row, list of the rest, and the
// device index of every receiver (for recording).  Positions follow
// pos_r = first + (spacing-1)*r (SRC/backward_modeling.c:108); when several
// receivers share a point the last one wins, as in the reference's loop.
static int upload_receivers(rtm_plan *q)
{
    const rtm_params &p = q->p;
    const int nrec = p.number_of_receivers;
    const int zr = p.thickness_up + p.depth_receiver;
    std::vector<int> recmap(q->X, -1), rec_j(nrec);
    std::map<long, int> outside;   // device index -> last receiver
    for (int r = 0; r < nrec; r++) {
        const int col = p.thickness_left + p.first_location_receiver + (p.spacing_receiver - 1) * r;
        if (col < 0 || col >= q->X || zr < 0 || zr >= q->Z) {
            rtm_set_error("receiver %d at padded (%d,%d) is outside the %dx%d grid", r, col, zr,
                          q->X, q->Z);
            return -1;
        }
        rec_j[r] = zr * q->P + col;
        const bool fused = col >= rtm::kHalf && col < q->X - rtm::kHalf && zr >= rtm::kHalf &&
                           zr < q->Z - rtm::kHalf;
        if (fused) recmap[col] = r;
        else outside[(long)zr * q->P + col] = r;
    }
    HIP_TRY(hipMemcpy(q->recmap, recmap.data(), sizeof(int) * q->X, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(q->rec_j, rec_j.data(), sizeof(int) * nrec, hipMemcpyHostToDevice));
    std::vector<int> oj, orr;
    for (auto &kv : outside) { oj.push_back((int)kv.first); orr.push_back(kv.second); }
    q->n_out = (int)oj.size();
    if (q->n_out) {
        HIP_TRY(hipMemcpy(q->out_j, oj.data(), sizeof(int) * q->n_out, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(q->out_r, orr.data(), sizeof(int) * q->n_out, hipMemcpyHostToDevice));
    }
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
    if (p->snapshot_mode != RTM_SNAP_FP32) {
        rtm_set_error("rtm_plan_create: snapshot mode %d is not available in this build", p->snapshot_mode);
        return -1;
    }
    HIP_TRY(hipSetDevice(p->device));
    rtm_plan *q = new rtm_plan();
    q->p = *p;
    q->real = real;
    q->esz = real == RTM_REAL_F64 ? 8 : 4;
    q->X = p->thickness_left + p->nx + p->thickness_right;
    q->Z = p->thickness_up + p->nz + p->thickness_down;
    q->P = round_up(q->X, 64);
    q->nfield = (size_t)q->P * q->Z;
    q->nint = (size_t)p->nx * p->nz;
    q->nsnap = (size_t)(p->number_of_time_steps / p->time_slices);
    const size_t nrec = (size_t)p->number_of_receivers, nt = (size_t)p->number_of_time_steps;
    auto fail = [&](void) { plan_free(q); delete q; return -1; };
#define ALLOC(ptr, bytes)                                                                \
    do {                                                                                 \
        hipError_t e_ = hipMalloc((void **)&(ptr), (bytes) ? (bytes) : 4);              \
        if (e_ != hipSuccess) {                                                          \
            rtm_set_error("hipMalloc(%zu) failed: %s", (size_t)(bytes), hipGetErrorString(e_)); \
            return fail();                                                               \
        }                                                                                \
    } while (0)
    if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&q->ev0) != hipSuccess || hipEventCreate(&q->ev1) != hipSuccess) {
        rtm_set_error("rtm_plan_create: stream/event creation failed");
        return fail();
    }
    const size_t fb = q->nfield * q->esz;
    ALLOC(q->u[0], fb); ALLOC(q->u[1], fb); ALLOC(q->C, fb);
    ALLOC(q->psiX[0], fb); ALLOC(q->psiX[1], fb); ALLOC(q->psiZ[0], fb); ALLOC(q->psiZ[1], fb);
    ALLOC(q->csiX, fb); ALLOC(q->csiZ, fb);
    ALLOC(q->ax, q->X * q->esz); ALLOC(q->bx, q->X * q->esz);
    ALLOC(q->az, q->Z * q->esz); ALLOC(q->bz, q->Z * q->esz);
    ALLOC(q->recmap, sizeof(int) * q->X);
    ALLOC(q->rec_j, sizeof(int) * nrec);
    ALLOC(q->out_j, sizeof(int) * nrec);
    ALLOC(q->out_r, sizeof(int) * nrec);
    ALLOC(q->wav, sizeof(float) * (size_t)p->sampling_wavelet);
    ALLOC(q->seis_t, sizeof(float) * nt * nrec);
    ALLOC(q->seis_rec, sizeof(float) * nt * nrec);
    ALLOC(q->snaps, q->nsnap * q->nint * q->esz);
    ALLOC(q->I, q->nint * q->esz); ALLOC(q->Is, q->nint * q->esz);
    ALLOC(q->Ic, q->nint * q->esz); ALLOC(q->Isc, q->nint * q->esz);
#undef ALLOC
    int rc = real == RTM_REAL_F64 ? upload_profiles<double>(q) : upload_profiles<float>(q);
    if (rc == 0) rc = upload_receivers(q);
    if (rc) return fail();
    if (hipMemset(q->I, 0, q->nint * q->esz) != hipSuccess ||
        hipMemset(q->Is, 0, q->nint * q->esz) != hipSuccess ||
        hipMemset(q->Ic, 0, q->nint * q->esz) != hipSuccess ||
        hipMemset(q->Isc, 0, q->nint * q->esz) != hipSuccess ||
        hipMemset(q->seis_rec, 0, sizeof(float) * nt * nrec) != hipSuccess) {
        rtm_set_error("rtm_plan_create: hipMemset failed");
        return fail();
    }
    *out = q;
    return 0;
}

extern "C" int rtm_plan_destroy(rtm_plan *q)
{
    if (!q) return 0;
    (void)hipSetDevice(q->p.device);
    (void)hipStreamSynchronize(q->stream);
    plan_free(q);
    delete q;
    return 0;
}

template <typename R>
static int set_velocity_t(rtm_plan *q, const float *vpad)
{
    std::vector<float> v(q->nfield);
    to_device_layout(vpad, v.data(), q->X, q->Z, q->P);
    float *dv = nullptr;
    HIP_TRY(hipMalloc(&dv, sizeof(float) * q->nfield));
    hipError_t e = hipMemcpyAsync(dv, v.data(), sizeof(float) * q->nfield, hipMemcpyHostToDevice, q->stream);
    if (e == hipSuccess) {
        // aux = dt/h rounded in R, as SRC/forward_modeling.c:71
        const R aux = sizeof(R) == 4 ? R(q->p.sampling / q->p.spacing_grid)
                                     : R((double)q->p.sampling / (double)q->p.spacing_grid);
        dim3 grid((q->X + 255) / 256, q->Z);
        hipLaunchKernelGGL(k_coef<R>, grid, dim3(256), 0, q->stream, dv, (R *)q->C, aux, q->X, q->Z, q->P);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(q->stream);
    (void)hipFree(dv);
    if (e != hipSuccess) { rtm_set_error("set_velocity: %s", hipGetErrorString(e)); return -1; }
    q->have_velocity = true;
    return 0;
}

extern "C" int rtm_plan_set_velocity(rtm_plan *q, const float *vpad)
{
    if (!q || !vpad) { rtm_set_error("rtm_plan_set_velocity: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->p.device));
    return q->real == RTM_REAL_F64 ? set_velocity_t<double>(q, vpad) : set_velocity_t<float>(q, vpad);
}

extern "C" int rtm_plan_set_wavelet(rtm_plan *q, const float *w, int n)
{
    if (!q || !w) { rtm_set_error("rtm_plan_set_wavelet: null argument"); return -1; }
    if (n != q->p.sampling_wavelet) {
        rtm_set_error("rtm_plan_set_wavelet: %d samples given, sampling_wavelet is %d", n,
                      q->p.sampling_wavelet);
        return -1;
    }
    q->wav_host.assign(w, w + n);
    HIP_TRY(hipSetDevice(q->p.device));
    HIP_TRY(hipMemcpy(q->wav, w, sizeof(float) * n, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int rtm_plan_set_seismogram(rtm_plan *q, const float *d)
{
    if (!q || !d) { rtm_set_error("rtm_plan_set_seismogram: null argument"); return -1; }
    const int nrec = q->p.number_of_receivers, nt = q->p.number_of_time_steps;
    std::vector<float> t((size_t)nt * nrec);
    for (int r = 0; r < nrec; r++)
        for (int n = 0; n < nt; n++) t[(size_t)n * nrec + r] = d[(size_t)r * nt + n];
    HIP_TRY(hipSetDevice(q->p.device));
    HIP_TRY(hipMemcpy(q->seis_t, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
    q->have_seis = true;
    return 0;
}

extern "C" int rtm_plan_zero_image(rtm_plan *q)
{
    if (!q) { rtm_set_error("rtm_plan_zero_image: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->p.device));
    const size_t b = q->nint * q->esz;
    HIP_TRY(hipMemsetAsync(q->I, 0, b, q->stream));
    HIP_TRY(hipMemsetAsync(q->Is, 0, b, q->stream));
    HIP_TRY(hipMemsetAsync(q->Ic, 0, b, q->stream));
    HIP_TRY(hipMemsetAsync(q->Isc, 0, b, q->stream));
    HIP_TRY(hipStreamSynchronize(q->stream));
    return 0;
}

// zero wavefields and CPML state: every modeling pass starts from rest
// (calloc at SRC/forward_modeling.c:43-67)
static int reset_state(rtm_plan *q)
{
    const size_t fb = q->nfield * q->esz;
    void *f[] = {q->u[0], q->u[1], q->psiX[0], q->psiX[1], q->psiZ[0], q->psiZ[1], q->csiX, q->csiZ};
    for (void *ptr : f) HIP_TRY(hipMemsetAsync(ptr, 0, fb, q->stream));
    q->cur = 0;
    q->psi_cur = 0;
    return 0;
}

template <typename R>
static StepArgs<R> base_args(rtm_plan *q)
{
    StepArgs<R> A{};
    const rtm_params &p = q->p;
    A.X = q->X; A.Z = q->Z; A.P = q->P;
    A.TL = p.thickness_left; A.TR = p.thickness_right; A.TU = p.thickness_up; A.TD = p.thickness_down;
    A.nx = p.nx; A.nz = p.nz; A.nrec = p.number_of_receivers;
    // h1 = 1/h, h2 = h*h, h3 = 1/(h*h), rounded as SRC/absorbingBoundaryCondition.c:159-161
    if (sizeof(R) == 4) {
        const float h = p.spacing_grid;
        A.h1 = R(1.0f / h); A.h2 = R(h * h); A.h3 = R(1.0f / (h * h));
    } else {
        const double h = p.spacing_grid;
        A.h1 = R(1.0 / h); A.h2 = R(h * h); A.h3 = R(1.0 / (h * h));
    }
    A.C = (const R *)q->C;
    A.ax = (const R *)q->ax; A.bx = (const R *)q->bx; A.az = (const R *)q->az; A.bz = (const R *)q->bz;
    A.csiX = (R *)q->csiX; A.csiZ = (R *)q->csiZ;
    A.src_j = -1;
    A.rec_row = p.thickness_up + p.depth_receiver;
    A.recmap = q->recmap;
    A.I = (R *)q->I; A.Is = (R *)q->Is; A.Ic = (R *)q->Ic; A.Isc = (R *)q->Isc;
    return A;
}

template <typename R>
static void bind_step(rtm_plan *q, StepArgs<R> &A)
{
    A.u = (const R *)q->u[q->cur];
    A.U = (R *)q->u[q->cur ^ 1];
    A.psiXi = (const R *)q->psiX[q->psi_cur];
    A.psiXo = (R *)q->psiX[q->psi_cur ^ 1];
    A.psiZi = (const R *)q->psiZ[q->psi_cur];
    A.psiZo = (R *)q->psiZ[q->psi_cur ^ 1];
}

static dim3 step_grid(const rtm_plan *q) { return dim3((q->X + kBX - 1) / kBX, (q->Z + kBZ - 1) / kBZ); }

// Forward modeling loop, SRC/forward_modeling.c:98-130.
template <typename R>
static int forward_t(rtm_plan *q, int shot_x, int shot_z, int record)
{
    const rtm_params &p = q->p;
    const int xs = shot_x + p.thickness_left, zs = shot_z + p.thickness_up;
    if (xs < 0 || xs >= q->X || zs < 0 || zs >= q->Z) {
        rtm_set_error("source at padded (%d,%d) is outside the %dx%d grid", xs, zs, q->X, q->Z);
        return -1;
    }
    const bool src_fused = xs >= rtm::kHalf && xs < q->X - rtm::kHalf && zs >= rtm::kHalf &&
                           zs < q->Z - rtm::kHalf;
    const int src_j = zs * q->P + xs;
    if (reset_state(q)) return -1;
    StepArgs<R> A = base_args<R>(q);
    const dim3 grid = step_grid(q), block(kBX, kBZ);
    const int TS = p.time_slices;
    HIP_TRY(hipEventRecord(q->ev0, q->stream));
    for (int n = 0; n < p.number_of_time_steps; n++) {
        bind_step(q, A);
        A.mode = rtm::M_STENCIL | rtm::M_CPML;
        const bool src = n < p.sampling_wavelet;
        if (src && src_fused) {
            A.mode |= rtm::M_SOURCE;
            A.src_j = src_j;
            A.src_w = R(q->wav_host[n]);
        }
        if (n % TS == 0) {
            A.mode |= rtm::M_SNAP;
            A.snap = (R *)q->snaps + (size_t)(n / TS) * q->nint;
        }
        hipLaunchKernelGGL(k_step<R>, grid, block, 0, q->stream, A);
        if (!src_fused) {
            // a source outside the stepped region can make the outer columns
            // non-zero, so their top-strip correction is applied too
            launch_top_edges(A, q->stream);
            if (src)
                hipLaunchKernelGGL(k_source_point<R>, dim3(1), dim3(64), 0, q->stream,
                                   (R *)q->u[q->cur ^ 1], (const R *)q->C, src_j, R(q->wav_host[n]));
        }
        if (record)
            hipLaunchKernelGGL(k_record<R>, dim3((p.number_of_receivers + 255) / 256), dim3(256), 0,
                               q->stream, (const R *)q->u[q->cur ^ 1], q->rec_j, p.number_of_receivers,
                               q->seis_rec + (size_t)n * p.number_of_receivers);
        q->cur ^= 1;
        q->psi_cur ^= 1;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(q->ev1, q->stream));
    HIP_TRY(hipEventSynchronize(q->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, q->ev0, q->ev1));
    q->fwd_ms = ms;
    q->last_steps = p.number_of_time_steps;
    q->have_snaps = true;
    return 0;
}

// Backward modeling + imaging loop, SRC/backward_modeling.c:97-126.
template <typename R>
static int backward_t(rtm_plan *q)
{
    const rtm_params &p = q->p;
    if (!q->have_snaps) { rtm_set_error("rtm_plan_backward: no snapshots (run forward first)"); return -1; }
    if (!q->have_seis) { rtm_set_error("rtm_plan_backward: no seismogram"); return -1; }
    if (reset_state(q)) return -1;
    StepArgs<R> A = base_args<R>(q);
    const dim3 grid = step_grid(q), block(kBX, kBZ);
    const int TS = p.time_slices, nt = p.number_of_time_steps, nrec = p.number_of_receivers;
    HIP_TRY(hipEventRecord(q->ev0, q->stream));
    for (int n = 0; n < nt; n++) {
        const int tau = nt - n - 1;
        bind_step(q, A);
        A.mode = rtm::M_STENCIL | rtm::M_CPML | rtm::M_INJECT;
        A.seis_row = q->seis_t + (size_t)tau * nrec;
        // snapshot tau/TS pairs with this step; the reference's counter
        // (SRC/backward_modeling.c:96,119) equals tau/TS when nt % TS == 0
        if (tau % TS == 0) {
            A.mode |= rtm::M_IMAGE | (p.image_kahan ? rtm::M_KAHAN : 0);
            A.snap = (R *)q->snaps + (size_t)(tau / TS) * q->nint;
        }
        hipLaunchKernelGGL(k_step<R>, grid, block, 0, q->stream, A);
        if (q->n_out) {
            // receivers outside the stepped region: outer-column top-strip
            // correction, then their replacement.  None of these points is in
            // the image region, so the fused imaging above is unaffected.
            launch_top_edges(A, q->stream);
            hipLaunchKernelGGL(k_inject_list<R>, dim3((q->n_out + 255) / 256), dim3(256), 0, q->stream,
                               (R *)q->u[q->cur ^ 1], q->out_j, q->out_r, q->n_out, A.seis_row);
        }
        q->cur ^= 1;
        q->psi_cur ^= 1;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(q->ev1, q->stream));
    HIP_TRY(hipEventSynchronize(q->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, q->ev0, q->ev1));
    q->bwd_ms = ms;
    q->last_steps = nt;
    return 0;
}

static int ready(rtm_plan *q, const char *who)
{
    if (!q) { rtm_set_error("%s: null plan", who); return -1; }
    if (!q->have_velocity) { rtm_set_error("%s: velocity not set", who); return -1; }
    if (hipSetDevice(q->p.device) != hipSuccess) { rtm_set_error("%s: hipSetDevice failed", who); return -1; }
    return 0;
}

extern "C" int rtm_plan_forward(rtm_plan *q, int shot_x, int shot_z, int record)
{
    if (ready(q, "rtm_plan_forward")) return -1;
    if (q->wav_host.size() != (size_t)q->p.sampling_wavelet) {
        rtm_set_error("rtm_plan_forward: wavelet not set");
        return -1;
    }
    return q->real == RTM_REAL_F64 ? forward_t<double>(q, shot_x, shot_z, record)
                                   : forward_t<float>(q, shot_x, shot_z, record);
}

extern "C" int rtm_plan_backward(rtm_plan *q)
{
    if (ready(q, "rtm_plan_backward")) return -1;
    return q->real == RTM_REAL_F64 ? backward_t<double>(q) : backward_t<float>(q);
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
    const int nrec = q->p.number_of_receivers, nt = q->p.number_of_time_steps;
    std::vector<float> t((size_t)nt * nrec);
    HIP_TRY(hipSetDevice(q->p.device));
    HIP_TRY(hipStreamSynchronize(q->stream));
    HIP_TRY(hipMemcpy(t.data(), q->seis_rec, sizeof(float) * t.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < nrec; r++)
        for (int n = 0; n < nt; n++) d[(size_t)r * nt + n] = t[(size_t)n * nrec + r];
    return 0;
}

template <typename R, typename O>
static int get_image_t(rtm_plan *q, O *I, O *Is)
{
    const int nx = q->p.nx, nz = q->p.nz;
    std::vector<R> h(q->nint);
    HIP_TRY(hipStreamSynchronize(q->stream));
    void *src[2] = {q->I, q->Is};
    O *dst[2] = {I, Is};
    for (int t = 0; t < 2; t++) {
        if (!dst[t]) continue;
        HIP_TRY(hipMemcpy(h.data(), src[t], sizeof(R) * q->nint, hipMemcpyDeviceToHost));
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) dst[t][(size_t)i * nz + k] = O(h[(size_t)k * nx + i]);
    }
    return 0;
}

extern "C" int rtm_plan_get_image(rtm_plan *q, float *I, float *Is)
{
    if (!q || !I) { rtm_set_error("rtm_plan_get_image: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->p.device));
    return q->real == RTM_REAL_F64 ? get_image_t<double, float>(q, I, Is) : get_image_t<float, float>(q, I, Is);
}

extern "C" int rtm_plan_get_image_f64(rtm_plan *q, double *I, double *Is)
{
    if (!q || !I) { rtm_set_error("rtm_plan_get_image_f64: null argument"); return -1; }
    HIP_TRY(hipSetDevice(q->p.device));
    return q->real == RTM_REAL_F64 ? get_image_t<double, double>(q, I, Is) : get_image_t<float, double>(q, I, Is);
}

template <typename R, typename O>
static int get_fields_t(rtm_plan *q, O *u_prev, O *u_last)
{
    std::vector<R> h(q->nfield);
    HIP_TRY(hipStreamSynchronize(q->stream));
    // after a loop u[cur] holds the last field and u[cur^1] the one before
    void *src[2] = {q->u[q->cur ^ 1], q->u[q->cur]};
    O *dst[2] = {u_prev, u_last};
    for (int t = 0; t < 2; t++) {
        if (!dst[t]) continue;
        HIP_TRY(hipMemcpy(h.data(), src[t], sizeof(R) * q->nfield, hipMemcpyDeviceToHost));
        from_device_layout(h.data(), dst[t], q->X, q->Z, q->P);
    }
    return 0;
}

extern "C" int rtm_plan_get_fields(rtm_plan *q, float *u_prev, float *u_last)
{
    if (!q) { rtm_set_error("rtm_plan_get_fields: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->p.device));
    return q->real == RTM_REAL_F64 ? get_fields_t<double, float>(q, u_prev, u_last)
                                   : get_fields_t<float, float>(q, u_prev, u_last);
}

extern "C" int rtm_plan_get_fields_f64(rtm_plan *q, double *u_prev, double *u_last)
{
    if (!q) { rtm_set_error("rtm_plan_get_fields_f64: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->p.device));
    return q->real == RTM_REAL_F64 ? get_fields_t<double, double>(q, u_prev, u_last)
                                   : get_fields_t<float, double>(q, u_prev, u_last);
}

extern "C" int rtm_plan_synchronize(rtm_plan *q)
{
    if (!q) { rtm_set_error("rtm_plan_synchronize: null plan"); return -1; }
    HIP_TRY(hipSetDevice(q->p.device));
    HIP_TRY(hipStreamSynchronize(q->stream));
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
