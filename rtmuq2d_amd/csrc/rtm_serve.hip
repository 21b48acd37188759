// rtm_serve.hip -- the request server: continuous batching of RTM requests of
// any grid size on one device.
//
// A request is one migration (a velocity realization, a wavelet, a
// seismogram and a shot position on its own grid).  Requests queue in
// arrival order; a tick admits every queued request whose HBM footprint
// (wavefields, CPML state, snapshot store, image) fits the device arena while
// a slot is free, then advances every active request by one block of time
// steps in the same launches: the tile and column lists of all active
// requests, whatever their grids, form one ragged work list, and every
// request carries its own phase and step counter, so requests in their
// forward pass, in their backward pass and just admitted share each launch.
// Between blocks, requests whose forward pass ended switch to their backward
// pass, and finished ones hand back their image and free their memory for
// the queue.  Each request's arithmetic is exactly that of a plan running it
// alone (the same kernels on the same per-request state), so its image does
// not depend on what else is in flight.
//
// The active requests advance in launch groups: greedily, in slot order,
// groups whose per-step working sets (16 B per padded grid point: u^n,
// u^{n-1}, C and the update) sum to at most group_bytes.  Each group has its
// own executor (work lists, captured graph) and runs its block of steps
// before the next group runs its own, so each launch's working set can stay
// in the device's 256 MB last-level cache.  Grouping changes only which
// requests share a launch, never a request's arithmetic.
//
// The reference runs one simulation per MPI rank, each start to finish
// (SRC/mpi_rtm_iso2d.c:39-51); this is the same work admitted as it arrives.
#include <deque>
#include <iterator>
#include <memory>

#include "rtm_sim.h"

using namespace rtmd;

namespace {

struct Request {
    int id = -1;
    rtm_params p{};
    std::vector<float> vpad, wav, seis;
    int sx = 0, sz = 0;
    SimBuffers b;
    int slot = -1;
    int phase = rtm::PH_IDLE;    // forward or backward once admitted
    int n = 0;                   // steps done in the current pass
    std::vector<float> image;    // [Nx][Nz] once complete
};

}  // namespace

struct rtm_server {
    int device = 0;
    int snapshot_mode = RTM_SNAP_FP32;
    int max_slots = 0, block = 100;
    int rz = 4, interior = rtm::IN_TILES, seg_rows = 0;
    DeviceArena arena;
    size_t group_bytes = kGroupBytes;   // launch-group budget (0: one group)
    std::vector<std::unique_ptr<Exec>> ex;      // one executor per launch group
    hipStream_t stream = nullptr;
    int next_id = 0;
    std::deque<std::unique_ptr<Request>> queue;
    std::vector<std::unique_ptr<Request>> slots;
    std::deque<std::unique_ptr<Request>> done;
    std::vector<SimDesc<float>> descs;
    bool lists_dirty = true, descs_dirty = true;
    long long steps_run = 0;
};

extern "C" size_t rtm_server_request_footprint(const rtm_params *p)
{
    if (!p || rtm_params_check(p)) return 0;
    // every buffer is rounded to the arena's 256-byte unit; ~40 buffers
    return sim_bytes(*p, 4) + 40 * 256;
}

static void server_free(rtm_server *s)
{
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &r : s->slots)
        if (r) r->b.release();
    s->slots.clear();
    for (auto &e : s->ex) e->release();
    s->ex.clear();
    s->arena.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->stream = nullptr;
}

extern "C" int rtm_server_create(int device, size_t hbm_bytes, int max_slots, int snapshot_mode, rtm_server **out)
{
    if (!out || max_slots < 1 || hbm_bytes < (1u << 20)) { rtm_set_error("rtm_server_create: bad argument"); return -1; }
    *out = nullptr;
    if (snapshot_mode < RTM_SNAP_FP32 || snapshot_mode > RTM_SNAP_BF16) {
        rtm_set_error("rtm_server_create: snapshot mode %d not available", snapshot_mode);
        return -1;
    }
    HIP_TRY(hipSetDevice(device));
    rtm_server *s = new rtm_server();
    s->device = device;
    s->snapshot_mode = snapshot_mode;
    s->max_slots = max_slots;
    s->slots.resize(max_slots);
    s->descs.assign(max_slots, SimDesc<float>{});
    auto fail = [&]() { server_free(s); delete s; return -1; };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        rtm_set_error("rtm_server_create: stream creation failed");
        return fail();
    }
    if (s->arena.init(hbm_bytes)) return fail();
    *out = s;
    return 0;
}

extern "C" int rtm_server_destroy(rtm_server *s)
{
    if (!s) return 0;
    (void)hipSetDevice(s->device);
    server_free(s);
    delete s;
    return 0;
}

extern "C" int rtm_server_set_tiling(rtm_server *s, int rows_per_thread, int graph_block, int interior, int seg_rows)
{
    if (!s || graph_block < 1) { rtm_set_error("rtm_server_set_tiling: bad argument"); return -1; }
    if (check_shape(rows_per_thread, interior)) return -1;
    s->rz = rows_per_thread;
    s->block = graph_block;
    s->interior = interior;
    s->seg_rows = seg_rows;
    s->lists_dirty = true;
    return 0;
}

extern "C" int rtm_server_submit(rtm_server *s, const rtm_params *p, const float *vpad, const float *wavelet,
                                 const float *seis, int shot_x, int shot_z)
{
    if (!s || !p || !vpad || !wavelet || !seis) { rtm_set_error("rtm_server_submit: null argument"); return -1; }
    if (rtm_params_check(p)) return -1;
    if (p->snapshot_mode != s->snapshot_mode) {
        rtm_set_error("rtm_server_submit: snapshot mode %d differs from the server's %d", p->snapshot_mode,
                      s->snapshot_mode);
        return -1;
    }
    const int X = p->thickness_left + p->nx + p->thickness_right;
    const int Z = p->thickness_up + p->nz + p->thickness_down;
    const int xs = shot_x + p->thickness_left, zs = shot_z + p->thickness_up;
    if (xs < rtm::kHalf || xs >= X - rtm::kHalf || zs < rtm::kHalf || zs >= Z - rtm::kHalf) {
        rtm_set_error("rtm_server_submit: source at padded (%d,%d) is outside the stepped region", xs, zs);
        return -1;
    }
    const int zr = p->thickness_up + p->depth_receiver;
    for (int r = 0; r < p->number_of_receivers; r++) {
        const int col = p->thickness_left + p->first_location_receiver + (p->spacing_receiver - 1) * r;
        if (col < rtm::kHalf || col >= X - rtm::kHalf || zr < rtm::kHalf || zr >= Z - rtm::kHalf) {
            rtm_set_error("rtm_server_submit: receiver %d at padded (%d,%d) is outside the stepped region", r, col, zr);
            return -1;
        }
    }
    if (rtm_server_request_footprint(p) > s->arena.size) {
        rtm_set_error("rtm_server_submit: request needs %zu bytes, the arena holds %zu", rtm_server_request_footprint(p),
                      s->arena.size);
        return -1;
    }
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->p = *p;
    r->vpad.assign(vpad, vpad + (size_t)X * Z);
    r->wav.assign(wavelet, wavelet + p->sampling_wavelet);
    r->seis.assign(seis, seis + (size_t)p->number_of_receivers * p->number_of_time_steps);
    r->sx = shot_x;
    r->sz = shot_z;
    const int id = r->id;
    s->queue.push_back(std::move(r));
    return id;
}

static SimDesc<float> forward_desc(const Request &r)
{
    SimDesc<float> d = r.b.desc<float>();
    d.phase = rtm::PH_FORWARD;
    d.flags = rtm::F_SNAP;
    d.src_j = (r.sz + r.p.thickness_up) * r.b.P + r.sx + r.p.thickness_left;
    return d;
}

static SimDesc<float> backward_desc(const Request &r)
{
    SimDesc<float> d = r.b.desc<float>();
    d.phase = rtm::PH_BACKWARD;
    d.flags = rtm::F_IMAGE | (r.p.image_kahan ? rtm::F_KAHAN : 0);
    return d;
}

// admit queued requests in arrival order while a slot is free and their
// buffers fit the arena; returns the number admitted (< 0 on error)
static int server_admit(rtm_server *s)
{
    int admitted = 0;
    while (!s->queue.empty()) {
        int slot = -1;
        for (int i = 0; i < s->max_slots && slot < 0; i++)
            if (!s->slots[i]) slot = i;
        if (slot < 0) break;
        Request &r = *s->queue.front();
        if (rtm_server_request_footprint(&r.p) > s->arena.size - s->arena.in_use) break;
        r.b.arena = &s->arena;
        if (r.b.alloc(r.p, RTM_REAL_F32)) {
            r.b.release();
            bool any = false;
            for (auto &a : s->slots) any = any || a;
            if (!any) {
                // nothing holds arena memory, so waiting cannot help: drop the
                // request rather than stall the queue behind it
                const int id = r.id;
                s->queue.pop_front();
                rtm_set_error("rtm_server_tick: request %d does not fit the empty arena; it was dropped", id);
                return -1;
            }
            rtm_set_error("");   // fragmentation: wait for memory to free up
            break;
        }
        // the host inputs stay with the request until it completes, so the
        // uploads need not wait
        if (r.b.set_velocity(r.vpad.data(), s->stream, false) || r.b.set_wavelet(r.wav.data(), (int)r.wav.size()) ||
            r.b.set_seismogram(r.seis.data(), s->stream, false))
            return -1;
        r.b.reset_state(s->stream);
        r.b.reset_snap_scales(s->stream);
        r.b.zero_image(s->stream);
        r.slot = slot;
        r.phase = rtm::PH_FORWARD;
        r.n = 0;
        s->descs[slot] = forward_desc(r);
        s->slots[slot] = std::move(s->queue.front());
        s->queue.pop_front();
        s->lists_dirty = s->descs_dirty = true;
        admitted++;
    }
    return admitted;
}

static size_t step_bytes(const Request &r) { return r.b.nfield * kStepBytesPerPoint; }

static int server_rebuild(rtm_server *s)
{
    // the last tick's graph launches may still run and read the executors'
    // lists and graphs: an executor is released or re-initialized below only
    // after they end (a graph destroyed in flight, or lists freed under
    // running kernels, is undefined)
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<std::vector<int>> groups;
    size_t sum = 0;
    for (int i = 0; i < s->max_slots; i++) {
        if (!s->slots[i]) continue;
        const size_t w = step_bytes(*s->slots[i]);
        if (groups.empty() || (s->group_bytes && sum + w > s->group_bytes)) {
            groups.emplace_back();
            sum = 0;
        }
        groups.back().push_back(i);
        sum += w;
    }
    while (s->ex.size() > groups.size()) {
        s->ex.back()->release();
        s->ex.pop_back();
    }
    while (s->ex.size() < groups.size()) s->ex.emplace_back(new Exec());
    for (size_t g = 0; g < groups.size(); g++) {
        std::vector<int4> tiles, inner;
        for (int i : groups[g]) append_work(tiles, inner, i, s->slots[i]->p, s->rz, s->interior, s->seg_rows);
        balance_tiles_for_xcds(tiles, 4 * s->rz);
        Exec &ex = *s->ex[g];
        if (!ex.d_desc || (int)tiles.size() > ex.cap_tiles || (int)inner.size() > ex.cap_inner || ex.rz != s->rz ||
            ex.interior != s->interior) {
            ex.release();
            const int ct = (int)(tiles.size() * 3 / 2) + 64, ci = (int)(inner.size() * 3 / 2) + 64;
            if (ex.init(RTM_REAL_F32, s->max_slots, ct, s->rz, ci, s->interior)) return -1;
            ex.q16 = s->snapshot_mode != RTM_SNAP_FP32;
        }
        if (ex.set_tiles(tiles, s->stream) || ex.set_inner(inner, s->stream)) return -1;
    }
    s->lists_dirty = false;
    s->descs_dirty = true;   // every executor gets the descriptors and counters of its new members
    return 0;
}

// the image of a finished request to the host ([Nx][Nz]); its memory back
static int server_finish(rtm_server *s, int slot)
{
    std::unique_ptr<Request> r = std::move(s->slots[slot]);
    const int nx = r->p.nx, nz = r->p.nz;
    std::vector<float> h(r->b.nint);
    HIP_TRY(hipMemcpyAsync(h.data(), r->b.I, sizeof(float) * r->b.nint, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    r->image.resize((size_t)nx * nz);
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) r->image[(size_t)i * nz + k] = h[(size_t)k * nx + i];
    r->b.release();
    r->vpad.clear(); r->vpad.shrink_to_fit();
    r->seis.clear(); r->seis.shrink_to_fit();
    r->slot = -1;
    s->descs[slot] = SimDesc<float>{};
    s->done.push_back(std::move(r));
    s->lists_dirty = s->descs_dirty = true;
    return 0;
}

extern "C" int rtm_server_tick(rtm_server *s)
{
    if (!s) { rtm_set_error("rtm_server_tick: null server"); return -1; }
    HIP_TRY(hipSetDevice(s->device));
    if (server_admit(s) < 0) return -1;
    int active = 0;
    for (auto &r : s->slots) active += r ? 1 : 0;
    if (active) {
        if (s->lists_dirty && server_rebuild(s)) return -1;
        if (s->descs_dirty) {
            // a request's step counter is the steps done in its current pass
            std::vector<int> n0(s->max_slots, 0);
            for (int i = 0; i < s->max_slots; i++)
                if (s->slots[i]) n0[i] = s->slots[i]->n;
            for (auto &e : s->ex)
                if (e->upload_descs(s->descs.data(), s->max_slots, s->stream) || e->set_n0(n0, s->stream)) return -1;
            s->descs_dirty = false;
        }
        // group after group, each a block of steps over its own members
        for (auto &e : s->ex)
            if (e->run(s->block, s->block, s->stream)) return -1;
        s->steps_run += s->block;
        // phase changes and completions at the block boundary
        for (int i = 0; i < s->max_slots; i++) {
            Request *r = s->slots[i].get();
            if (!r) continue;
            r->n += s->block;
            if (r->n < r->p.number_of_time_steps) continue;
            if (r->phase == rtm::PH_FORWARD) {
                if (r->b.flush_last_snapshot(s->stream)) return -1;
                r->b.reset_state(s->stream);
                r->phase = rtm::PH_BACKWARD;
                r->n = 0;
                s->descs[i] = backward_desc(*r);
                s->descs_dirty = true;
            } else if (server_finish(s, i)) {
                return -1;
            }
        }
        HIP_TRY(hipGetLastError());
    }
    int left = (int)s->queue.size();
    for (auto &r : s->slots) left += r ? 1 : 0;
    return left;
}

extern "C" int rtm_server_next_done(rtm_server *s, int *id)
{
    if (!s || !id) { rtm_set_error("rtm_server_next_done: null argument"); return -1; }
    if (s->done.empty()) return 0;
    *id = s->done.front()->id;
    return 1;
}

extern "C" int rtm_server_take_image(rtm_server *s, int id, float *image, size_t n)
{
    if (!s || !image) { rtm_set_error("rtm_server_take_image: null argument"); return -1; }
    for (auto it = s->done.begin(); it != s->done.end(); ++it) {
        if ((*it)->id != id) continue;
        if (n != (*it)->image.size()) {
            rtm_set_error("rtm_server_take_image: request %d has %zu image points, not %zu", id, (*it)->image.size(), n);
            return -1;
        }
        std::copy((*it)->image.begin(), (*it)->image.end(), image);
        s->done.erase(it);
        return 0;
    }
    rtm_set_error("rtm_server_take_image: request %d is not complete", id);
    return -1;
}

extern "C" int rtm_server_set_group_bytes(rtm_server *s, size_t bytes)
{
    if (!s) { rtm_set_error("rtm_server_set_group_bytes: null server"); return -1; }
    s->group_bytes = bytes;
    s->lists_dirty = true;
    return 0;
}

extern "C" int rtm_server_get_group_bytes(rtm_server *s, size_t *bytes, int *groups)
{
    if (!s) { rtm_set_error("rtm_server_get_group_bytes: null server"); return -1; }
    if (bytes) *bytes = s->group_bytes;
    if (groups) *groups = (int)s->ex.size();
    return 0;
}

extern "C" int rtm_server_stats(rtm_server *s, int *active, int *queued, size_t *bytes_in_use, long long *steps)
{
    if (!s) { rtm_set_error("rtm_server_stats: null server"); return -1; }
    int a = 0;
    for (auto &r : s->slots) a += r ? 1 : 0;
    if (active) *active = a;
    if (queued) *queued = (int)s->queue.size();
    if (bytes_in_use) *bytes_in_use = s->arena.in_use;
    if (steps) *steps = s->steps_run;
    return 0;
}
