// rtm_sim.hip -- per-simulation device buffers and the tiled step executor.
#include <mutex>


#include "rtm_sim.h"

namespace rtmd {

int DeviceArena::init(size_t bytes)
{
    release();
    size = rounded(bytes);
    HIP_TRY(hipMalloc((void **)&base, size));
    free_[0] = size;
    return 0;
}

void DeviceArena::release()
{
    if (base) (void)hipFree(base);
    base = nullptr;
    size = in_use = 0;
    free_.clear();
    live_.clear();
}

void *DeviceArena::alloc(size_t bytes)
{
    const size_t need = rounded(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < need) continue;
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > need) free_[off + need] = len - need;
        live_[off] = need;
        in_use += need;
        return base + off;
    }
    return nullptr;
}

void DeviceArena::free(void *ptr)
{
    if (!ptr || !base) return;
    const size_t off = (size_t)((char *)ptr - base);
    auto lv = live_.find(off);
    if (lv == live_.end()) return;
    size_t start = off, len = lv->second;
    in_use -= len;
    live_.erase(lv);
    auto nx = free_.lower_bound(start);
    if (nx != free_.end() && start + len == nx->first) {   // merge with the next block
        len += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {                            // merge with the previous block
        auto pv = std::prev(nx);
        if (pv->first + pv->second == start) {
            start = pv->first;
            len += pv->second;
            free_.erase(pv);
        }
    }
    free_[start] = len;
}

size_t DeviceArena::largest_free() const
{
    size_t m = 0;
    for (const auto &f : free_) m = f.second > m ? f.second : m;
    return m;
}

size_t sim_bytes(const rtm_params &p, size_t esz)
{
    const int X = p.thickness_left + p.nx + p.thickness_right;
    const int Z = p.thickness_up + p.nz + p.thickness_down;
    const size_t P = (size_t)round_up(X, 64);
    const size_t nfield = P * Z;
    const size_t nint = (size_t)p.nx * p.nz;
    const size_t nsnap = (size_t)(p.number_of_time_steps / p.time_slices);
    const size_t nrec = (size_t)p.number_of_receivers, nt = (size_t)p.number_of_time_steps;
    size_t b = 0;
    b += 9 * nfield * esz;                          // u x2, C, psiX x2, psiZ x2, csiX, csiZ
    b += (size_t)(2 * X + 2 * Z) * esz;             // profiles
    b += (size_t)X * 4 + 3 * nrec * 4;              // receiver tables
    b += (size_t)p.sampling_wavelet * 4;
    b += 2 * nt * nrec * 4;                         // input and recorded traces
    if (p.snapshot_mode == RTM_SNAP_FP16 || p.snapshot_mode == RTM_SNAP_BF16)
        b += nsnap * nint * 2 + nsnap * rtm::kSnapWords * 4;   // 16-bit snapshot store + max slots
    else if (p.snapshot_mode == RTM_SNAP_I16)
        b += nsnap * nint * 2 + nsnap * p.nz * rtm::i16_blocks(p.thickness_left, p.nx);   // codes + block exponents
    else if (p.snapshot_mode == RTM_SNAP_BOUNDARY)
        b += 2 * nfield * esz + nt * (size_t)(8 * (p.nx + 8) + 8 * p.nz) * esz;   // source field + ring saves
    else
        b += nsnap * nint * esz;                    // snapshot store
    b += 4 * nint * esz;                            // I, Is and their compensations
    return b;
}

int SimBuffers::alloc(const rtm_params &pp, int real_type, const SimBuffers *share)
{
    p = pp;
    real = real_type;
    esz = real == RTM_REAL_F64 ? 8 : 4;
    X = p.thickness_left + p.nx + p.thickness_right;
    Z = p.thickness_up + p.nz + p.thickness_down;
    P = round_up(X, 64);
    nfield = (size_t)P * Z;
    nint = (size_t)p.nx * p.nz;
    nsnap = (size_t)(p.number_of_time_steps / p.time_slices);
    const size_t nrec = (size_t)p.number_of_receivers, nt = (size_t)p.number_of_time_steps;
    auto get = [&](void **ptr, size_t bytes) -> int {
        if (arena) {
            *ptr = arena->alloc(bytes ? bytes : 4);
            if (!*ptr) {
                rtm_set_error("device arena: no free block of %zu bytes", bytes);
                return -1;
            }
        } else {
            hipError_t e = hipMalloc(ptr, bytes ? bytes : 4);
            if (e != hipSuccess) {
                rtm_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
                return -1;
            }
        }
        owned.push_back(*ptr);
        return 0;
    };
    const size_t fb = nfield * esz;
    int rc = 0;
    rc |= get(&u[0], fb); rc |= rc ? 0 : get(&u[1], fb);
    if (!rc) rc = get(&C, fb);
    if (!rc) rc = get(&psiX[0], fb);
    if (!rc) rc = get(&psiX[1], fb);
    if (!rc) rc = get(&psiZ[0], fb);
    if (!rc) rc = get(&psiZ[1], fb);
    if (!rc) rc = get(&csiX, fb);
    if (!rc) rc = get(&csiZ, fb);
    qmode = p.snapshot_mode == RTM_SNAP_FP16 ? rtm::Q_FP16
            : p.snapshot_mode == RTM_SNAP_BF16 ? rtm::Q_BF16
            : p.snapshot_mode == RTM_SNAP_I16  ? rtm::Q_I16 : rtm::Q_NONE;
    qb0 = rtm::i16_first_block(p.thickness_left);
    nqb = rtm::i16_blocks(p.thickness_left, p.nx);
    ringmode = p.snapshot_mode == RTM_SNAP_BOUNDARY;
    NR = 8 * (p.nx + 8) + 8 * p.nz;
    if (!rc && qmode != rtm::Q_NONE && real == RTM_REAL_F64) {
        rtm_set_error("16-bit snapshot stores are for FP32 stepping (the FP64 plan is the oracle run)");
        rc = -1;
    }
    if (!rc && ringmode && (p.thickness_left < rtm::kHalf || p.thickness_right < rtm::kHalf ||
                            p.thickness_up < rtm::kHalf || p.thickness_down < rtm::kHalf)) {
        rtm_set_error("boundary saving needs absorbing layers at least %d points thick (the saved ring lies in them)",
                      rtm::kHalf);
        rc = -1;
    }
    if (ringmode) {
        if (!rc) rc = get(&sbuf[0], fb);
        if (!rc) rc = get(&sbuf[1], fb);
        if (!rc) rc = get(&ring, nt * (size_t)NR * esz);
    } else if (qmode == rtm::Q_NONE) {
        if (!rc) rc = get(&snaps, nsnap * nint * esz);
    } else if (qmode == rtm::Q_I16) {
        if (!rc) rc = get((void **)&q16, nsnap * nint * 2);
        if (!rc) rc = get((void **)&qexp, nsnap * (size_t)p.nz * nqb);
    } else {
        if (!rc) rc = get((void **)&q16, nsnap * nint * 2);
        if (!rc) rc = get((void **)&smax, nsnap * rtm::kSnapWords * 4);
    }
    if (!rc) rc = get(&I, nint * esz);
    if (!rc) rc = get(&Is, nint * esz);
    if (!rc) rc = get(&Ic, nint * esz);
    if (!rc) rc = get(&Isc, nint * esz);
    if (!rc && share) {
        ax = share->ax; bx = share->bx; az = share->az; bz = share->bz;
        recmap = share->recmap; rec_j = share->rec_j; out_j = share->out_j; out_r = share->out_r;
        n_out = share->n_out; rec_dup = share->rec_dup;
        wav = share->wav; seis_t = share->seis_t; seis_rec = share->seis_rec;
        wav_host = share->wav_host;
        have_seis = share->have_seis;
    } else if (!rc) {
        if (!rc) rc = get(&ax, X * esz);
        if (!rc) rc = get(&bx, X * esz);
        if (!rc) rc = get(&az, Z * esz);
        if (!rc) rc = get(&bz, Z * esz);
        if (!rc) rc = get((void **)&recmap, sizeof(int) * X);
        if (!rc) rc = get((void **)&rec_j, sizeof(int) * nrec);
        if (!rc) rc = get((void **)&out_j, sizeof(int) * nrec);
        if (!rc) rc = get((void **)&out_r, sizeof(int) * nrec);
        if (!rc) rc = get((void **)&wav, sizeof(float) * (size_t)p.sampling_wavelet);
        if (!rc) rc = get((void **)&seis_t, sizeof(float) * nt * nrec);
        if (!rc) rc = get((void **)&seis_rec, sizeof(float) * nt * nrec);
        if (!rc) rc = upload_profiles();
        if (!rc) rc = upload_receivers();
        if (!rc && (hipMemset(seis_rec, 0, sizeof(float) * nt * nrec) != hipSuccess ||
                    hipMemset(seis_t, 0, sizeof(float) * nt * nrec) != hipSuccess)) {
            rtm_set_error("SimBuffers: hipMemset failed");
            rc = -1;
        }
    }
    if (!rc) {
        // velocity-independent state starts at zero
        void *z[] = {I, Is, Ic, Isc, C};
        for (void *ptr : z)
            if (hipMemset(ptr, 0, ptr == C ? fb : nint * esz) != hipSuccess) rc = -1;
        if (rc) rtm_set_error("SimBuffers: hipMemset failed");
    }
    if (rc) release();
    return rc;
}

void SimBuffers::release()
{
    for (void *ptr : owned) {
        if (arena) arena->free(ptr);
        else (void)hipFree(ptr);
    }
    owned.clear();
    u[0] = u[1] = u[2] = u[3] = C = psiX[0] = psiX[1] = psiZ[0] = psiZ[1] = csiX = csiZ = nullptr;
    umask = 1;
    ax = bx = az = bz = nullptr;
    recmap = rec_j = out_j = out_r = nullptr;
    wav = seis_t = seis_rec = nullptr;
    snaps = I = Is = Ic = Isc = nullptr;
    q16 = nullptr;
    smax = nullptr;
    qexp = nullptr;
    sbuf[0] = sbuf[1] = ring = nullptr;
}

template <typename R>
static int upload_profiles_t(SimBuffers &q)
{
    const rtm_params &p = q.p;
    const int TL = p.thickness_left, TR = p.thickness_right, TU = p.thickness_up, TD = p.thickness_down;
    std::vector<R> aL(TL), aR(TR), aT(TU), aB(TD), bL(TL), bR(TR), bT(TU), bB(TD);
    if constexpr (sizeof(R) == 4)
        rtm_profiles_f32(&p, aL.data(), aR.data(), aT.data(), aB.data(), bL.data(), bR.data(), bT.data(), bB.data());
    else
        rtm_profiles_f64(&p, aL.data(), aR.data(), aT.data(), aB.data(), bL.data(), bR.data(), bT.data(), bB.data());
    // full-length profiles: x strips use the left/right arrays, z strips the
    // top/bottom arrays, indexed as SRC/absorbingBoundaryCondition.c:181,273,227,320
    std::vector<R> ax(q.X, R(0)), bx(q.X, R(0)), az(q.Z, R(0)), bz(q.Z, R(0));
    for (int i = 0; i < TL; i++) { ax[i] = aL[i]; bx[i] = bL[i]; }
    for (int i = 0; i < TR; i++) { ax[q.X - TR + i] = aR[i]; bx[q.X - TR + i] = bR[i]; }
    for (int k = 0; k < TU; k++) { az[k] = aT[k]; bz[k] = bT[k]; }
    for (int k = 0; k < TD; k++) { az[q.Z - TD + k] = aB[k]; bz[q.Z - TD + k] = bB[k]; }
    HIP_TRY(hipMemcpy(q.ax, ax.data(), sizeof(R) * q.X, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(q.bx, bx.data(), sizeof(R) * q.X, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(q.az, az.data(), sizeof(R) * q.Z, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(q.bz, bz.data(), sizeof(R) * q.Z, hipMemcpyHostToDevice));
    return 0;
}

int SimBuffers::upload_profiles()
{
    return real == RTM_REAL_F64 ? upload_profiles_t<double>(*this) : upload_profiles_t<float>(*this);
}

// Receiver tables.  Positions follow pos_r = first + (spacing-1)*r
// (SRC/backward_modeling.c:108); when several receivers share a point the
// last one wins, as in the reference's loop.
int SimBuffers::upload_receivers()
{
    const int nrec = p.number_of_receivers;
    const int zr = p.thickness_up + p.depth_receiver;
    std::vector<int> rm(X, -1), rj(nrec);
    std::map<long, int> outside;
    rec_dup = false;
    for (int r = 0; r < nrec; r++) {
        const int col = p.thickness_left + p.first_location_receiver + (p.spacing_receiver - 1) * r;
        if (col < 0 || col >= X || zr < 0 || zr >= Z) {
            rtm_set_error("receiver %d at padded (%d,%d) is outside the %dx%d grid", r, col, zr, X, Z);
            return -1;
        }
        rj[r] = zr * P + col;
        const bool inside = col >= rtm::kHalf && col < X - rtm::kHalf && zr >= rtm::kHalf && zr < Z - rtm::kHalf;
        if (inside) {
            if (rm[col] >= 0) rec_dup = true;
            rm[col] = r;
        } else {
            if (outside.count((long)zr * P + col)) rec_dup = true;
            outside[(long)zr * P + col] = r;
        }
    }
    HIP_TRY(hipMemcpy(recmap, rm.data(), sizeof(int) * X, hipMemcpyHostToDevice));
    if (nrec) HIP_TRY(hipMemcpy(rec_j, rj.data(), sizeof(int) * nrec, hipMemcpyHostToDevice));
    std::vector<int> oj, orr;
    for (auto &kv : outside) { oj.push_back((int)kv.first); orr.push_back(kv.second); }
    n_out = (int)oj.size();
    if (n_out) {
        HIP_TRY(hipMemcpy(out_j, oj.data(), sizeof(int) * n_out, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(out_r, orr.data(), sizeof(int) * n_out, hipMemcpyHostToDevice));
    }
    return 0;
}

// The padded velocity arrives in the reference's layout ([X][Z], z fastest)
// and is staged in u[0] (X*Z floats fit in the field, and every pass zeroes
// u before stepping); k_coef reads it transposed.  No allocation and no
// host-side transposition or wait, so requests can be admitted while others
// run.  The wavefields read back before the next pass are overwritten.
template <typename R>
static int set_velocity_t(SimBuffers &q, const float *vpad, hipStream_t s, bool wait)
{
    float *stage = (float *)q.u[0];
    HIP_TRY(hipMemcpyAsync(stage, vpad, sizeof(float) * (size_t)q.X * q.Z, hipMemcpyHostToDevice, s));
    // aux = dt/h rounded in R, as SRC/forward_modeling.c:71
    const R aux = sizeof(R) == 4 ? R(q.p.sampling / q.p.spacing_grid)
                                 : R((double)q.p.sampling / (double)q.p.spacing_grid);
    hipLaunchKernelGGL(k_coef<R>, dim3((q.X + 255) / 256, q.Z), dim3(256), 0, s, stage, (R *)q.C, aux, q.X, q.Z,
                       q.P);
    HIP_TRY(hipGetLastError());
    if (wait) HIP_TRY(hipStreamSynchronize(s));
    q.have_velocity = true;
    return 0;
}

int SimBuffers::set_velocity(const float *vpad, hipStream_t s, bool wait)
{
    return real == RTM_REAL_F64 ? set_velocity_t<double>(*this, vpad, s, wait)
                                : set_velocity_t<float>(*this, vpad, s, wait);
}

int SimBuffers::set_wavelet(const float *w, int n)
{
    if (n != p.sampling_wavelet) {
        rtm_set_error("wavelet: %d samples given, sampling_wavelet is %d", n, p.sampling_wavelet);
        return -1;
    }
    wav_host.assign(w, w + n);
    if (n) HIP_TRY(hipMemcpy(wav, w, sizeof(float) * n, hipMemcpyHostToDevice));
    return 0;
}

// [nrec][nt] -> [nt][nrec] through 32 x 32 LDS tiles
__global__ void k_transpose(const float *__restrict__ in, float *__restrict__ out, int rows, int cols)
{
    __shared__ float t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int r = r0 + j, c = c0 + threadIdx.x;
        if (r < rows && c < cols) t[j][threadIdx.x] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int c = c0 + j, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = t[threadIdx.x][j];
    }
}

// The traces arrive [nrec][nt] and are staged in the recording buffer (the
// forward pass's recorded traces are overwritten), then transposed on the
// device: no host-side transposition or wait.
int SimBuffers::set_seismogram(const float *d, hipStream_t s, bool wait)
{
    const int nrec = p.number_of_receivers, nt = p.number_of_time_steps;
    HIP_TRY(hipMemcpyAsync(seis_rec, d, sizeof(float) * (size_t)nrec * nt, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_transpose, dim3((nt + 31) / 32, (nrec + 31) / 32), dim3(32, 8), 0, s, seis_rec, seis_t,
                       nrec, nt);
    HIP_TRY(hipGetLastError());
    if (wait) HIP_TRY(hipStreamSynchronize(s));
    have_seis = true;
    return 0;
}

int SimBuffers::set_quad(bool on)
{
    if (on && !u[3]) {
        const size_t fb = nfield * esz;
        for (int t = 2; t < 4; t++) {
            if (arena) {
                u[t] = arena->alloc(fb);
                if (!u[t]) {
                    rtm_set_error("device arena: no free block of %zu bytes", fb);
                    return -1;
                }
            } else {
                hipError_t e = hipMalloc(&u[t], fb);
                if (e != hipSuccess) {
                    u[t] = nullptr;
                    rtm_set_error("hipMalloc(%zu) failed: %s", fb, hipGetErrorString(e));
                    return -1;
                }
            }
            owned.push_back(u[t]);
        }
    }
    umask = on ? 3 : 1;
    return 0;
}

void SimBuffers::reset_state(hipStream_t s, bool source_field)
{
    void *w0 = source_field ? sbuf[0] : u[0], *w1 = source_field ? sbuf[1] : u[1];
    void *w2 = !source_field && umask == 3 ? u[2] : nullptr, *w3 = !source_field && umask == 3 ? u[3] : nullptr;
    if (real == RTM_REAL_F64) {
        double *f[] = {(double *)w0, (double *)w1, (double *)w2, (double *)w3, (double *)psiX[0],
                       (double *)psiX[1], (double *)psiZ[0], (double *)psiZ[1], (double *)csiX, (double *)csiZ};
        for (double *ptr : f) fill_zero(ptr, nfield, s);
    } else {
        float *f[] = {(float *)w0, (float *)w1, (float *)w2, (float *)w3, (float *)psiX[0],
                      (float *)psiX[1], (float *)psiZ[0], (float *)psiZ[1], (float *)csiX, (float *)csiZ};
        for (float *ptr : f) fill_zero(ptr, nfield, s);
    }
}

void SimBuffers::reset_snap_scales(hipStream_t s)
{
    if (smax) fill_zero(smax, nsnap * rtm::kSnapWords, s);
}

// The 16-bit store is filled one step late, from u^n of the step after the
// snapshot step (rtm::step_view); a snapshot taken on the pass's last step
// (TS = 1) is filled here from the final field instead.
template <typename R>
__global__ void k_quantize_interior(const R *__restrict__ u, int P, int TL, int TU, int nx, int nz,
                                    unsigned short *__restrict__ q, const unsigned *__restrict__ slots, int mode)
{
    // every lane of the wavefront takes part in reading the slots
    const float scale = rtm::pow2f(14 - rtm::snap_exp(rtm::snap_max_bits(slots)));
    const int ii = blockIdx.x * blockDim.x + threadIdx.x, kk = blockIdx.y;
    if (ii >= nx || kk >= nz) return;
    q[(size_t)kk * nx + ii] = rtm::quant16(float(u[(size_t)(kk + TU) * P + ii + TL]), scale, mode);
}

int SimBuffers::flush_last_snapshot(hipStream_t s)
{
    const int nt = p.number_of_time_steps, TS = p.time_slices;
    // the int16 block store writes each snapshot in the step that computes it
    if (qmode == rtm::Q_NONE || qmode == rtm::Q_I16 || (nt - 1) % TS != 0) return 0;
    const int c = (nt - 1) / TS;
    const dim3 grid((p.nx + 255) / 256, p.nz);
    hipLaunchKernelGGL(k_quantize_interior<float>, grid, dim3(256), 0, s, (const float *)u[nt & umask], P,
                       p.thickness_left, p.thickness_up, p.nx, p.nz, q16 + (size_t)c * nint,
                       smax + (size_t)c * rtm::kSnapWords, qmode);
    HIP_TRY(hipGetLastError());
    return 0;
}

void SimBuffers::zero_image(hipStream_t s)
{
    if (real == RTM_REAL_F64) {
        double *f[] = {(double *)I, (double *)Is, (double *)Ic, (double *)Isc};
        for (double *ptr : f) fill_zero(ptr, nint, s);
    } else {
        float *f[] = {(float *)I, (float *)Is, (float *)Ic, (float *)Isc};
        for (float *ptr : f) fill_zero(ptr, nint, s);
    }
}

template <typename R>
SimDesc<R> SimBuffers::desc() const
{
    SimDesc<R> d{};
    d.X = X; d.Z = Z; d.P = P;
    d.TL = p.thickness_left; d.TR = p.thickness_right; d.TU = p.thickness_up; d.TD = p.thickness_down;
    d.nx = p.nx; d.nz = p.nz;
    d.nt = p.number_of_time_steps; d.TS = p.time_slices; d.nw = p.sampling_wavelet;
    d.nrec = p.number_of_receivers;
    d.phase = rtm::PH_IDLE;
    d.flags = 0;
    d.src_j = -1;
    d.rec_row = p.thickness_up + p.depth_receiver;
    // h1 = 1/h, h2 = h*h, h3 = 1/(h*h), rounded as SRC/absorbingBoundaryCondition.c:159-161
    if (sizeof(R) == 4) {
        const float h = p.spacing_grid;
        d.h1 = R(1.0f / h); d.h2 = R(h * h); d.h3 = R(1.0f / (h * h));
    } else {
        const double h = p.spacing_grid;
        d.h1 = R(1.0 / h); d.h2 = R(h * h); d.h3 = R(1.0 / (h * h));
    }
    for (int t = 0; t < 4; t++) d.ubuf[t] = (R *)u[t];
    d.umask = umask;
    d.C = (const R *)C;
    d.ax = (const R *)ax; d.bx = (const R *)bx; d.az = (const R *)az; d.bz = (const R *)bz;
    d.psiX[0] = (R *)psiX[0]; d.psiX[1] = (R *)psiX[1];
    d.psiZ[0] = (R *)psiZ[0]; d.psiZ[1] = (R *)psiZ[1];
    d.csiX = (R *)csiX; d.csiZ = (R *)csiZ;
    d.recmap = recmap;
    d.wav = wav;
    d.seis_t = seis_t;
    d.seis_rec = seis_rec;
    d.snaps = (R *)snaps;
    d.qmode = qmode;
    d.q16 = q16;
    d.smax = smax;
    d.qexp = qexp;
    d.qb0 = qb0;
    d.nqb = nqb;
    d.sbuf[0] = (R *)sbuf[0];
    d.sbuf[1] = (R *)sbuf[1];
    d.ring = (R *)ring;
    d.NR = NR;
    d.I = (R *)I; d.Is = (R *)Is; d.Ic = (R *)Ic; d.Isc = (R *)Isc;
    return d;
}

template SimDesc<float> SimBuffers::desc<float>() const;
template SimDesc<double> SimBuffers::desc<double>() const;

void append_work(std::vector<int4> &tiles, std::vector<int4> &inner, int sim, const rtm_params &p, int rz,
                 int interior, int seg_rows)
{
    std::vector<rtm::WorkItem> t, in;
    rtm::build_work(t, in, sim, p.thickness_left + p.nx + p.thickness_right,
                    p.thickness_up + p.nz + p.thickness_down, p.thickness_left, p.thickness_right, p.thickness_up,
                    p.thickness_down, rz, interior, seg_rows);
    for (const auto &w : t) tiles.push_back(make_int4(w.a, w.b, w.c, w.d));
    for (const auto &w : in) inner.push_back(make_int4(w.a, w.b, w.c, w.d));
}

void balance_tiles_for_xcds(std::vector<int4> &tiles, int tz)
{
    std::vector<rtm::WorkItem> w(tiles.size());
    for (size_t i = 0; i < tiles.size(); i++) w[i] = {tiles[i].x, tiles[i].y, tiles[i].z, tiles[i].w};
    rtm::balance_for_xcds(w, tz);
    tiles.resize(w.size());
    for (size_t i = 0; i < w.size(); i++) tiles[i] = make_int4(w[i].a, w[i].b, w[i].c, w[i].d);
}

// ---------------------------------------------------------------------------
// executor
// ---------------------------------------------------------------------------

int check_shape(int rz, int interior)
{
    if (rz != 1 && rz != 2 && rz != 4 && rz != 8) {
        rtm_set_error("tile rows per thread must be 1, 2, 4 or 8 (got %d)", rz);
        return -1;
    }
    if (interior != rtm::IN_TILES && interior != rtm::IN_COLUMNS && interior != rtm::IN_MIXED &&
        interior != rtm::IN_TILES8 && interior != rtm::IN_PAIRS) {
        rtm_set_error("interior kernel must be 0 (tiles), 1 (columns), 2 (mixed tiles), 3 (tiles of 8 "
                      "wavefronts) or 4 (two-step pair items) (got %d)", interior);
        return -1;
    }
    if ((interior == rtm::IN_MIXED || interior == rtm::IN_TILES8) && rz != 2 && rz != 4) {
        rtm_set_error("%s take 2 or 4 rows per thread (got %d)",
                      interior == rtm::IN_MIXED ? "mixed tiles" : "tiles of 8 wavefronts", rz);
        return -1;
    }
    return 0;
}

int Exec::init(int real_type, int max_sims, int max_tiles, int rows_per_thread, int max_inner, int interior_kernel)
{
    real = real_type;
    rz = rows_per_thread;
    interior = interior_kernel;
    if (check_shape(rz, interior)) return -1;
    cap_sims = max_sims > 0 ? max_sims : 1;
    cap_tiles = max_tiles > 0 ? max_tiles : 1;
    const size_t dsz = real == RTM_REAL_F64 ? sizeof(SimDesc<double>) : sizeof(SimDesc<float>);
    HIP_TRY(hipMalloc(&d_desc, dsz * cap_sims));
    HIP_TRY(hipMalloc((void **)&d_n0, sizeof(int) * cap_sims));
    HIP_TRY(hipMalloc((void **)&d_tiles, sizeof(int4) * cap_tiles));
    HIP_TRY(hipMalloc((void **)&d_ntiles, sizeof(int)));
    HIP_TRY(hipMemset(d_desc, 0, dsz * cap_sims));
    HIP_TRY(hipMemset(d_n0, 0, sizeof(int) * cap_sims));
    HIP_TRY(hipMemset(d_ntiles, 0, sizeof(int)));
    cap_inner = max_inner > 0 ? max_inner : 1;
    HIP_TRY(hipMalloc((void **)&d_inner, sizeof(int4) * cap_inner));
    HIP_TRY(hipMalloc((void **)&d_ninner, sizeof(int)));
    HIP_TRY(hipMemset(d_ninner, 0, sizeof(int)));
    HIP_TRY(hipMalloc((void **)&d_err, sizeof(int)));
    HIP_TRY(hipMemset(d_err, 0, sizeof(int)));
    sched = false;
    ninner = 0;
    h_n0.assign(cap_sims, 0);
    // the side stream is made when an interior list first needs it
    // (set_inner): every stream of the process takes a share of the
    // device's few hardware queues, and idle ones only crowd them
    HIP_TRY(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    return 0;
}

static void destroy_captured(Exec::Captured &c)
{
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c = Exec::Captured{};
}

int Exec::drop_graph()
{
    bool any = false;
    for (const Captured &c : cap) any = any || c.exec;
    if (any && g_stream) (void)hipStreamSynchronize(g_stream);   // no graph is destroyed in flight
    for (Captured &c : cap) destroy_captured(c);
    return 0;
}

void Exec::release()
{
    drop_graph();
    void *all[] = {d_desc, d_n0, d_tiles, d_ntiles, d_inner, d_ninner, d_err};
    for (void *p : all)
        if (p) (void)hipFree(p);
    d_desc = nullptr; d_n0 = nullptr; d_tiles = nullptr; d_ntiles = nullptr;
    d_inner = nullptr; d_ninner = nullptr; d_err = nullptr;
    sched = false;
    ninner = ntiles = 0;
    if (side) {
        (void)hipStreamSynchronize(side);
        (void)hipStreamDestroy(side);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    side = nullptr;
    ev_fork = ev_join = nullptr;
}

int Exec::set_inner(const std::vector<int4> &items, hipStream_t s)
{
    if ((int)items.size() > cap_inner) {
        rtm_set_error("interior list of %zu exceeds the executor capacity %d", items.size(), cap_inner);
        return -1;
    }
    ninner = (int)items.size();
    if (ninner && !side && interior == rtm::IN_COLUMNS) {
        HIP_TRY(hipStreamSynchronize(s));
        drop_graph();   // graphs captured without the fork
        HIP_TRY(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    }
    if (ninner) HIP_TRY(hipMemcpyAsync(d_inner, items.data(), sizeof(int4) * ninner, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_ninner, &ninner, sizeof(int), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if ((ninner > 0) != g_inner) drop_graph();
    return 0;
}

int Exec::set_tiles(const std::vector<int4> &tiles, hipStream_t s)
{
    if ((int)tiles.size() > cap_tiles) {
        rtm_set_error("tile list of %zu exceeds the executor capacity %d", tiles.size(), cap_tiles);
        return -1;
    }
    ntiles = (int)tiles.size();
    if (ntiles) HIP_TRY(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int4) * ntiles, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_ntiles, &ntiles, sizeof(int), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if ((ntiles > 0) != g_tiles) drop_graph();
    return 0;
}

int Exec::upload_descs(const void *descs, int n, hipStream_t s)
{
    if (n > cap_sims) { rtm_set_error("%d simulations exceed the executor capacity %d", n, cap_sims); return -1; }
    nsims = n;
    const size_t dsz = real == RTM_REAL_F64 ? sizeof(SimDesc<double>) : sizeof(SimDesc<float>);
    HIP_TRY(hipMemcpyAsync(d_desc, descs, dsz * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int Exec::set_n0(const std::vector<int> &n0, hipStream_t s)
{
    const int n = (int)n0.size() < cap_sims ? (int)n0.size() : cap_sims;
    for (int i = 0; i < n; i++) h_n0[i] = n0[i];
    HIP_TRY(hipMemcpyAsync(d_n0, h_n0.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

// The light (KIND K_LIGHT: no snapshot store, no imaging) or the store
// (K_STORE: the snapshot store, no imaging; I16: the int16 block store)
// tile instance for FP32 fields in tile shapes; false when there is none.
template <typename R, bool Q16, int UM, int KIND, bool I16 = false>
static bool launch_kind(const Exec &e, int s_in_block, hipStream_t ts)
{
    if constexpr (Q16 || sizeof(R) != 4) {
        return false;
    } else {
        const SimDesc<R> *d = (const SimDesc<R> *)e.d_desc;
        const int grid = 8 * ((e.cap_tiles + 7) / 8);
        const bool one = e.cap_sims == 1;
        switch (e.rz) {
        case 1:
            if (one) hipLaunchKernelGGL((k_tiles<R, 1, false, true, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<R, 1, false, false, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            break;
        case 2:
            if (one) hipLaunchKernelGGL((k_tiles<R, 2, false, true, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<R, 2, false, false, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            break;
        case 8:
            if (one) hipLaunchKernelGGL((k_tiles<R, 8, false, true, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<R, 8, false, false, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            break;
        default:
            if (one) hipLaunchKernelGGL((k_tiles<R, 4, false, true, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<R, 4, false, false, I16, 4, UM, KIND>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            break;
        }
        return true;
    }
}

template <typename R, bool Q16>
static void launch_tiles(const Exec &e, int s_in_block, hipStream_t s, int kind)
{
    const SimDesc<R> *d = (const SimDesc<R> *)e.d_desc;
    // interior and strip tiles write disjoint points and read only u^n and
    // their own points: the tiles go to the side stream, joined at the end
    const bool fork = e.ninner && e.ntiles && e.side;
    hipStream_t ts = s;
    if (fork) {
        (void)hipEventRecord(e.ev_fork, s);
        (void)hipStreamWaitEvent(e.side, e.ev_fork, 0);
        ts = e.side;
    }
    if (e.ninner) {
        // four items per workgroup
        const int cgrid = 8 * (((e.cap_inner + 3) / 4 + 7) / 8);
        hipLaunchKernelGGL((k_columns<R, Q16>), dim3(cgrid), dim3(256), 0, s, d, e.d_n0, e.d_inner, e.d_ninner,
                           s_in_block);
    }
    if (e.ntiles && e.interior == rtm::IN_MIXED) {
        const int grid = 8 * ((e.cap_tiles + 7) / 8);
        if (e.rz == 2)
            hipLaunchKernelGGL((k_tiles_mix<R, 2, Q16>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block);
        else
            hipLaunchKernelGGL((k_tiles_mix<R, 4, Q16>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block);
    } else if (e.ntiles && e.interior == rtm::IN_TILES8) {
        const int grid = 8 * ((e.cap_tiles + 7) / 8);
        const bool one = e.cap_sims == 1;
        if (e.rz == 2) {
            if (one) hipLaunchKernelGGL((k_tiles<R, 2, Q16, true, false, 8>), dim3(grid), dim3(512), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<R, 2, Q16, false, false, 8>), dim3(grid), dim3(512), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
        } else {
            if (one) hipLaunchKernelGGL((k_tiles<R, 4, Q16, true, false, 8>), dim3(grid), dim3(512), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<R, 4, Q16, false, false, 8>), dim3(grid), dim3(512), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
        }
    } else if (e.ntiles && kind == K_LIGHT && launch_kind<R, Q16, 1, K_LIGHT>(e, s_in_block, ts)) {
        // a step that neither stores nor images: the light instance
    } else if (e.ntiles && kind == K_STORE && launch_kind<R, Q16, 1, K_STORE>(e, s_in_block, ts)) {
        // a forward step that stores a snapshot: the store instance
    } else if (e.ntiles && e.cap_sims == 1) {
        const int grid = 8 * ((e.cap_tiles + 7) / 8);
        switch (e.rz) {
        case 1: hipLaunchKernelGGL((k_tiles<R, 1, Q16, true>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 2: hipLaunchKernelGGL((k_tiles<R, 2, Q16, true>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 8: hipLaunchKernelGGL((k_tiles<R, 8, Q16, true>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        default: hipLaunchKernelGGL((k_tiles<R, 4, Q16, true>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        }
    } else if (e.ntiles) {
        const int grid = 8 * ((e.cap_tiles + 7) / 8);
        switch (e.rz) {
        case 1: hipLaunchKernelGGL((k_tiles<R, 1, Q16>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 2: hipLaunchKernelGGL((k_tiles<R, 2, Q16>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 8: hipLaunchKernelGGL((k_tiles<R, 8, Q16>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        default: hipLaunchKernelGGL((k_tiles<R, 4, Q16>), dim3(grid), dim3(256), 0, ts, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        }
    }
    if (fork) {
        (void)hipEventRecord(e.ev_join, e.side);
        (void)hipStreamWaitEvent(s, e.ev_join, 0);
    }
}

template <typename R>
static void launch_ring(const Exec &e, int s_in_block, hipStream_t s)
{
    const SimDesc<R> *d = (const SimDesc<R> *)e.d_desc;
    const dim3 grid(e.ring_blocks, e.cap_sims);
    if (e.ring_phase == rtm::PH_FORWARD)
        hipLaunchKernelGGL(k_ring_save<R>, grid, dim3(256), 0, s, d, e.d_n0, s_in_block);
    else if (e.ring_phase == rtm::PH_BACKWARD)
        hipLaunchKernelGGL(k_recon<R>, grid, dim3(256), 0, s, d, e.d_n0, s_in_block);
}

// the int16 block store's launch: tile lists only (check_store_shape)
static void launch_tiles_i16(const Exec &e, int s_in_block, hipStream_t s)
{
    if (!e.ntiles) return;
    const SimDesc<float> *d = (const SimDesc<float> *)e.d_desc;
    const int grid = 8 * ((e.cap_tiles + 7) / 8);
    if (e.interior == rtm::IN_TILES8) {
        const bool one = e.cap_sims == 1;
        if (e.rz == 2) {
            if (one) hipLaunchKernelGGL((k_tiles<float, 2, false, true, true, 8>), dim3(grid), dim3(512), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<float, 2, false, false, true, 8>), dim3(grid), dim3(512), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
        } else {
            if (one) hipLaunchKernelGGL((k_tiles<float, 4, false, true, true, 8>), dim3(grid), dim3(512), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
            else hipLaunchKernelGGL((k_tiles<float, 4, false, false, true, 8>), dim3(grid), dim3(512), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err);
        }
        return;
    }
    if (e.cap_sims == 1) {
        switch (e.rz) {
        case 1: hipLaunchKernelGGL((k_tiles<float, 1, false, true, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 2: hipLaunchKernelGGL((k_tiles<float, 2, false, true, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 8: hipLaunchKernelGGL((k_tiles<float, 8, false, true, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        default: hipLaunchKernelGGL((k_tiles<float, 4, false, true, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        }
    } else {
        switch (e.rz) {
        case 1: hipLaunchKernelGGL((k_tiles<float, 1, false, false, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 2: hipLaunchKernelGGL((k_tiles<float, 2, false, false, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 8: hipLaunchKernelGGL((k_tiles<float, 8, false, false, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        default: hipLaunchKernelGGL((k_tiles<float, 4, false, false, true>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        }
    }
}

int check_store_shape(int snapshot_mode, int interior)
{
    if (snapshot_mode == RTM_SNAP_I16 && interior != rtm::IN_TILES && interior != rtm::IN_TILES8) {
        rtm_set_error("the int16 block snapshot store runs in tile shapes only (interior kernel 0 or 3, got %d)",
                      interior);
        return -1;
    }
    if (interior == rtm::IN_PAIRS && snapshot_mode != RTM_SNAP_FP32) {
        rtm_set_error("two-step pair items (interior kernel 4) run with the FP32 snapshot store only");
        return -1;
    }
    return 0;
}

int check_pair_real(int real, int interior)
{
    if (interior == rtm::IN_PAIRS && real != RTM_REAL_F32) {
        rtm_set_error("two-step pair items (interior kernel 4) step FP32 fields only");
        return -1;
    }
    return 0;
}

// Launch of step s of a block with pair items: the pair items at s and s+1
// (an even s), then the tiles of the rest of the grid at s (k_tiles over
// four u buffers).  FP32, FP32 store (check_store_shape, check_pair_real).
static void launch_pairs(const Exec &e, int s_in_block, int lim, hipStream_t s, bool heavy)
{
    const SimDesc<float> *d = (const SimDesc<float> *)e.d_desc;
    const bool one = e.cap_sims == 1;
    if (e.ninner && (s_in_block & 1) == 0) {
        const int grid = 8 * (((e.cap_inner + 3) / 4 + 7) / 8);
        // prefetch distance 1: distances 2 and 3 measured the same on the c2
        // launch group (the march is not waiting on its loads)
        if (one) hipLaunchKernelGGL((k_pairs<float, true, 1>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_inner, e.d_ninner, s_in_block, lim);
        else hipLaunchKernelGGL((k_pairs<float, false, 1>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_inner, e.d_ninner, s_in_block, lim);
    }
    if (!e.ntiles) return;
    if (!heavy && launch_kind<float, false, 3, K_LIGHT>(e, s_in_block, s)) return;
    const int grid = 8 * ((e.cap_tiles + 7) / 8);
    if (one) {
        switch (e.rz) {
        case 1: hipLaunchKernelGGL((k_tiles<float, 1, false, true, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 2: hipLaunchKernelGGL((k_tiles<float, 2, false, true, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 8: hipLaunchKernelGGL((k_tiles<float, 8, false, true, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        default: hipLaunchKernelGGL((k_tiles<float, 4, false, true, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        }
    } else {
        switch (e.rz) {
        case 1: hipLaunchKernelGGL((k_tiles<float, 1, false, false, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 2: hipLaunchKernelGGL((k_tiles<float, 2, false, false, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        case 8: hipLaunchKernelGGL((k_tiles<float, 8, false, false, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        default: hipLaunchKernelGGL((k_tiles<float, 4, false, false, false, 4, 3>), dim3(grid), dim3(256), 0, s, d, e.d_n0, e.d_tiles, e.d_ntiles, s_in_block, e.d_err); break;
        }
    }
}

int Exec::launch_one(int s_in_block, hipStream_t s, int lim, int kind)
{
    if (interior == rtm::IN_PAIRS) launch_pairs(*this, s_in_block, lim, s, kind != K_LIGHT);
    else if (real == RTM_REAL_F64) launch_tiles<double, false>(*this, s_in_block, s, K_FULL);
    else if (i16) {
        // the light instance is the FP32 one (a light step stores nothing);
        // the store instance carries the int16 block store
        const bool done = interior == rtm::IN_TILES &&
                          (kind == K_LIGHT   ? launch_kind<float, false, 1, K_LIGHT>(*this, s_in_block, s)
                           : kind == K_STORE ? launch_kind<float, false, 1, K_STORE, true>(*this, s_in_block, s)
                                             : false);
        if (!done) launch_tiles_i16(*this, s_in_block, s);
    }
    else if (q16) launch_tiles<float, true>(*this, s_in_block, s, K_FULL);
    else launch_tiles<float, false>(*this, s_in_block, s, kind);
    // boundary saving: the ring save / reconstruction follows the step on the
    // same stream (running them beside the step on a stream of their own cost
    // more in graph edges than it hid)
    if (ring_phase && ring_blocks > 0) {
        if (real == RTM_REAL_F64) launch_ring<double>(*this, s_in_block, s);
        else launch_ring<float>(*this, s_in_block, s);
    }
    return 0;
}

void Exec::set_schedule(int phase, int nt, int ts, int flags)
{
    sched = phase == rtm::PH_FORWARD || phase == rtm::PH_BACKWARD;
    sched_phase = sched ? phase : 0;
    sched_nt = nt;
    sched_ts = ts > 0 ? ts : 1;
    sched_flags = flags;
    gsched.clear();
}

void Exec::set_group_schedule(const std::vector<int4> &per_sim)
{
    sched = false;
    gsched = per_sim;
    for (auto &g : gsched) g.z = g.z > 0 ? g.z : 1;
}

// the instance a step of a simulation's pass needs (StepKind): a forward
// snapshot step the store one, a backward imaging step the full one; the
// pattern is periodic in n (also past the pass's end, where the kernels
// skip the simulation), so a group's steps repeat from block to block
static int kind_of(const int4 &g, int n)
{
    const auto mod = [](int a, int m) { return ((a % m) + m) % m; };
    if (g.x == rtm::PH_FORWARD) return (g.w & rtm::F_SNAP) && mod(n, g.z) == 0 ? K_STORE : K_LIGHT;
    if (g.x == rtm::PH_BACKWARD) return (g.w & rtm::F_IMAGE) && mod(g.y - 1 - n, g.z) == 0 ? K_FULL : K_LIGHT;
    return K_LIGHT;
}

int Exec::heavy_at(int t) const
{
    if (sched) return heavy_step(h_n0[0] + t);
    if (gsched.empty()) return K_FULL;
    int k = K_LIGHT;
    for (int i = 0; i < (int)gsched.size() && i < cap_sims; i++) k = std::max(k, kind_of(gsched[i], h_n0[i] + t));
    return k;
}

bool Exec::pattern(int block) const
{
    if (sched) return block % sched_ts == 0;
    if (gsched.empty()) return false;
    for (const auto &g : gsched)
        if (g.x && block % g.z) return false;
    return true;
}

int Exec::heavy_step(int n) const
{
    if (!sched) return K_FULL;
    if (n < 0 || n >= sched_nt) return K_LIGHT;   // inactive: the kernels skip the simulation
    if (sched_phase == rtm::PH_FORWARD) return (sched_flags & rtm::F_SNAP) && n % sched_ts == 0 ? K_STORE : K_LIGHT;
    return (sched_flags & rtm::F_IMAGE) && (sched_nt - 1 - n) % sched_ts == 0 ? K_FULL : K_LIGHT;
}


int Exec::check_err(hipStream_t s)
{
    int h = 0;
    HIP_TRY(hipMemcpyAsync(&h, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!h) return 0;
    fill_zero(d_err, 1, s);
    (void)hipStreamSynchronize(s);
    rtm_set_error("internal: %s", (h & 2) ? "a step launch met a corner tile that the work list should have cut"
                                            : "a snapshot-store step launch met an imaging step (step schedule mismatch)");
    return -1;
}

int Exec::set_ring_phase(int phase, int blocks)
{
    ring_phase = phase;
    ring_blocks = phase ? blocks : 0;
    return 0;
}

int Exec::run(int steps, int block, hipStream_t s)
{
    if (steps <= 0) return 0;
    if (!d_desc || !d_tiles || !d_ntiles) {
        rtm_set_error("the step executor holds no work lists (its last tiling call failed)");
        return -1;
    }
    int done = 0;
    if (block > 1 && steps >= block) {
        // with a step schedule whose pattern repeats from block to block, the
        // captured steps run the light or the full tile instance as the
        // schedule says; otherwise every captured step runs the full one
        std::vector<char> heavy;
        if (pattern(block)) {
            heavy.resize(block);
            for (int t = 0; t < block; t++) heavy[t] = heavy_at(t);
        }
        int slot = sched ? sched_phase : ring_phase;
        if (!sched && !gsched.empty()) {
            slot = -1;
            for (int k = 3; k < 7 && slot < 0; k++)
                if (cap[k].exec && cap[k].block == block && cap[k].ring_blocks == ring_blocks && cap[k].heavy == heavy)
                    slot = k;
            if (slot < 0) slot = 3 + (gcap_next++ & 3);
        }
        Captured &c = cap[slot];
        const bool lists_match = g_tiles == (ntiles > 0) && g_inner == (ninner > 0);
        if (!lists_match) drop_graph();
        if (!c.exec || c.block != block || c.ring_blocks != ring_blocks || c.heavy != heavy) {
            if (c.exec && g_stream) (void)hipStreamSynchronize(g_stream);
            destroy_captured(c);
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int t = 0; t < block; t++) launch_one(t, s, block, heavy.empty() ? (int)K_FULL : (int)heavy[t]);
            hipLaunchKernelGGL(k_advance, dim3((cap_sims + 255) / 256), dim3(256), 0, s, d_n0, cap_sims, block);
            hipError_t e = hipStreamEndCapture(s, &c.graph);
            if (e != hipSuccess) { rtm_set_error("graph capture failed: %s", hipGetErrorString(e)); return -1; }
            HIP_TRY(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
            c.block = block;
            c.ring_blocks = ring_blocks;
            c.heavy = heavy;
            g_tiles = ntiles > 0;
            g_inner = ninner > 0;
        }
        g_stream = s;
        while (steps - done >= block) {
            HIP_TRY(hipGraphLaunch(c.exec, s));
            done += block;
        }
    }
    const int rem = steps - done;
    for (int t = 0; t < rem; t++) launch_one(t, s, rem, heavy_at(done + t));
    if (rem) hipLaunchKernelGGL(k_advance, dim3((cap_sims + 255) / 256), dim3(256), 0, s, d_n0, cap_sims, rem);
    HIP_TRY(hipGetLastError());
    for (int i = 0; i < cap_sims; i++) h_n0[i] += steps;
    return 0;
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------

const std::vector<Shape> &tuner_candidates()
{
    static const std::vector<Shape> c = {
        {1, rtm::IN_TILES, 0}, {2, rtm::IN_TILES, 0}, {4, rtm::IN_TILES, 0}, {8, rtm::IN_TILES, 0},
        {4, rtm::IN_COLUMNS, 0}, {4, rtm::IN_COLUMNS, 64}, {2, rtm::IN_COLUMNS, 32}, {1, rtm::IN_COLUMNS, 8},
        {4, rtm::IN_MIXED, 0}, {2, rtm::IN_MIXED, 0},
        {2, rtm::IN_TILES8, 0}, {4, rtm::IN_TILES8, 0},
        {2, rtm::IN_PAIRS, 24},
    };
    return c;
}

namespace {
std::mutex g_tune_mu;
std::map<std::string, Shape> g_tune;
}

std::string tuner_key(int real, int qmode, const std::vector<const rtm_params *> &grids)
{
    std::string k = std::to_string(real) + "q" + std::to_string(qmode);
    for (const rtm_params *p : grids) {
        char b[96];
        snprintf(b, sizeof b, "|%d,%d,%d,%d,%d,%d", p->nx, p->nz, p->thickness_left, p->thickness_right,
                 p->thickness_up, p->thickness_down);
        k += b;
    }
    return k;
}

bool tuner_lookup(const std::string &key, Shape *out)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    *out = it->second;
    return true;
}

void tuner_store(const std::string &key, const Shape &s)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = s;
}

template <typename R>
int time_steps(Exec &ex, std::vector<SimDesc<R>> descs, int steps, int block, hipStream_t s, double *ms, int flags)
{
    for (auto &d : descs) {
        d.phase = rtm::PH_FORWARD;
        d.flags = flags;
        d.src_j = -1;
    }
    if (ex.upload_descs(descs.data(), (int)descs.size(), s)) return -1;
    if (ex.set_n0(std::vector<int>(descs.size(), 0), s)) return -1;
    // the forward schedule of these steps, as a pass runs them
    ex.set_schedule(descs.empty() || (descs[0].qmode != rtm::Q_NONE && descs[0].qmode != rtm::Q_I16) ? 0
                                                                                                   : rtm::PH_FORWARD,
                    descs.empty() ? 0 : descs[0].nt, descs.empty() ? 1 : descs[0].TS, flags);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = ex.run(block > 1 ? block : 10, block, s);      // warm (and capture)
    if (!rc) rc = ex.set_n0(std::vector<int>(descs.size(), 0), s);
    if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = -1;
    if (!rc) rc = ex.run(steps, block, s);
    float t = 0.f;
    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess)) {
        rtm_set_error("time_steps: event failure");
        rc = -1;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc) rc = ex.check_err(s);
    *ms = t / (steps > 0 ? steps : 1);
    return rc;
}

template int time_steps<float>(Exec &, std::vector<SimDesc<float>>, int, int, hipStream_t, double *, int);
template int time_steps<double>(Exec &, std::vector<SimDesc<double>>, int, int, hipStream_t, double *, int);

}  // namespace rtmd
