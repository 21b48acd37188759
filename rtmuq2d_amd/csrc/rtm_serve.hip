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
// The queue, the slots, the groups and the failure rules are rtm_sched.h's
// (device-independent, tested on the CPU); this file is its HIP backend.  A
// failing launch group fails its own requests only, a failed admission fails
// that request alone, and submissions may come from any thread.
//
// The reference runs one simulation per MPI rank, each start to finish
// (SRC/mpi_rtm_iso2d.c:39-51); this is the same work admitted as it arrives.
#include <memory>

#include "rtm_sim.h"
#include "rtm_sched.h"

using namespace rtmd;

namespace {

struct ReqState {
    SimBuffers b;
};
using SReq = rtm_sched::Req<ReqState>;

SimDesc<float> forward_desc(const SReq &r)
{
    SimDesc<float> d = r.st.b.desc<float>();
    d.phase = rtm::PH_FORWARD;
    d.flags = rtm::F_SNAP;
    d.src_j = (r.sz + r.p.thickness_up) * r.st.b.P + r.sx + r.p.thickness_left;
    return d;
}

SimDesc<float> backward_desc(const SReq &r)
{
    SimDesc<float> d = r.st.b.desc<float>();
    d.phase = rtm::PH_BACKWARD;
    d.flags = rtm::F_IMAGE | (r.p.image_kahan ? rtm::F_KAHAN : 0);
    return d;
}

constexpr int kServerLanes = 4;

// The device side of the server: the arena, one executor per launch group,
// one stream, and up to three more lanes: with lanes > 1 launch group g steps
// on lane g % lanes, so the groups of a tick run at once and one group's
// steps fill the gaps between another's launches.  Every other call first
// joins the lanes into the stream (join), so nothing on the stream overtakes
// a group's steps.  Every call runs under the core's tick lock.
struct HipBackend {
    using State = ReqState;
    int device = 0;
    int snapshot_mode = RTM_SNAP_FP32;
    int max_slots = 0;
    int rz = 4, interior = rtm::IN_TILES, seg_rows = 0;
    DeviceArena arena;
    std::vector<std::unique_ptr<Exec>> ex;      // one executor per launch group
    std::vector<std::vector<int>> members_of;   // group g's request slots (as of its last prepare)
    hipStream_t stream = nullptr;
    int lanes = 1;
    hipStream_t lane_s[kServerLanes] = {};      // lane 0 is `stream`
    hipEvent_t lane_ev[kServerLanes] = {}, fork_ev = nullptr;
    bool lanes_busy = false;
    std::string err;

    // the stream waits for every lane's queued steps
    void join()
    {
        if (!lanes_busy) return;
        for (int j = 1; j < kServerLanes; j++)
            if (lane_s[j]) {
                (void)hipEventRecord(lane_ev[j], lane_s[j]);
                (void)hipStreamWaitEvent(stream, lane_ev[j], 0);
            }
        lanes_busy = false;
    }

    std::string error() const { return err; }
    int fail(const char *what)
    {
        err = std::string(what) + ": " + rtm_last_error();
        return -1;
    }
    int hip_fail(const char *what, hipError_t e)
    {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return -1;
    }

    int admit(SReq &r, int slot)
    {
        (void)slot;
        join();
        if (rtm_server_request_footprint(&r.p) > arena.size - arena.in_use) return 1;
        r.st.b.arena = &arena;
        if (r.st.b.alloc(r.p, RTM_REAL_F32)) {   // no block large enough: wait or fail
            r.st.b.release();
            return 1;
        }
        // the host inputs stay with the request until it completes, so the
        // uploads need not wait
        if (r.st.b.set_velocity(r.vpad.data(), stream, false)) return fail("velocity upload");
        if (r.st.b.set_wavelet(r.wav.data(), (int)r.wav.size())) return fail("wavelet upload");
        if (r.st.b.set_seismogram(r.seis.data(), stream, false)) return fail("seismogram upload");
        r.st.b.reset_state(stream);
        r.st.b.reset_snap_scales(stream);
        r.st.b.zero_image(stream);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail("admission", e);
        return 0;
    }

    void release(SReq &r)
    {
        join();
        r.st.b.release();
    }

    size_t step_bytes(const SReq &r) const { return r.st.b.nfield * kStepBytesPerPoint; }

    int prepare(int g, const std::vector<const SReq *> &members)
    {
        join();
        while ((int)ex.size() <= g) ex.emplace_back(new Exec());
        std::vector<int4> tiles, inner;
        for (const SReq *r : members) append_work(tiles, inner, r->slot, r->p, rz, interior, seg_rows);
        balance_tiles_for_xcds(tiles, rtm::tile_rows(rz, interior));
        Exec &e = *ex[g];
        if (!e.d_desc || (int)tiles.size() > e.cap_tiles || (int)inner.size() > e.cap_inner || e.rz != rz ||
            e.interior != interior) {
            e.release();
            const int ct = (int)(tiles.size() * 3 / 2) + 64, ci = (int)(inner.size() * 3 / 2) + 64;
            if (e.init(RTM_REAL_F32, max_slots, ct, rz, ci, interior)) return fail("executor");
            e.q16 = snapshot_mode == RTM_SNAP_FP16 || snapshot_mode == RTM_SNAP_BF16;
            e.i16 = snapshot_mode == RTM_SNAP_I16;
        }
        if (e.set_tiles(tiles, stream) || e.set_inner(inner, stream)) return fail("work lists");
        if ((int)members_of.size() <= g) members_of.resize(g + 1);
        members_of[g].clear();
        for (const SReq *r : members) members_of[g].push_back(r->slot);
        return 0;
    }

    // the caller has synchronised the stream: no launch reads these executors
    void trim(int ngroups)
    {
        join();
        while ((int)ex.size() > ngroups) {
            ex.back()->release();
            ex.pop_back();
        }
        if ((int)members_of.size() > ngroups) members_of.resize(ngroups);
    }

    void reset(int g)
    {
        join();
        if (g < (int)ex.size()) ex[g]->release();
        if (g < (int)members_of.size()) members_of[g].clear();
    }

    int upload(int g, const std::vector<const SReq *> &slots)
    {
        join();
        std::vector<SimDesc<float>> d(max_slots, SimDesc<float>{});
        std::vector<int> n0(max_slots, 0);   // a request's counter: steps done in its current pass
        for (int i = 0; i < max_slots; i++) {
            if (!slots[i]) continue;
            d[i] = slots[i]->pass == rtm_sched::P_FORWARD ? forward_desc(*slots[i]) : backward_desc(*slots[i]);
            n0[i] = slots[i]->n;
        }
        if (ex[g]->upload_descs(d.data(), max_slots, stream) || ex[g]->set_n0(n0, stream))
            return fail("descriptors");
        // the members' step schedules: a step at which none of them stores
        // or images runs the light tile instance (the FP32 and int16 block
        // stores write a snapshot only in the step that computes it)
        std::vector<int4> gs;
        if ((snapshot_mode == RTM_SNAP_FP32 || snapshot_mode == RTM_SNAP_I16) && g < (int)members_of.size()) {
            gs.assign(max_slots, make_int4(0, 0, 1, 0));
            for (int i : members_of[g])
                if (i >= 0 && i < max_slots && slots[i])
                    gs[i] = make_int4(d[i].phase, d[i].nt, d[i].TS, d[i].flags);
        }
        ex[g]->set_group_schedule(gs);
        return 0;
    }

    int run(int g, int steps)
    {
        (void)hipGetLastError();   // an earlier group's error is that group's
        const int j = lanes > 1 ? g % lanes : 0;
        hipStream_t s = stream;
        if (j > 0) {
            // the lane starts after everything queued on the stream so far
            hipError_t e = hipEventRecord(fork_ev, stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(lane_s[j], fork_ev, 0);
            if (e != hipSuccess) return hip_fail("lane", e);
            s = lane_s[j];
            lanes_busy = true;
        }
        if (ex[g]->run(steps, steps, s)) return fail("launch group");
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail("launch group", e);
        return 0;
    }

    int set_lanes(int n)
    {
        join();
        if (hipStreamSynchronize(stream) != hipSuccess) return fail("lanes");
        for (int j = 1; j < n; j++) {
            if (!lane_s[j] && hipStreamCreateWithFlags(&lane_s[j], hipStreamNonBlocking) != hipSuccess)
                return fail("lane stream");
            if (!lane_ev[j] && hipEventCreateWithFlags(&lane_ev[j], hipEventDisableTiming) != hipSuccess)
                return fail("lane event");
        }
        if (!fork_ev && hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming) != hipSuccess)
            return fail("lane event");
        lanes = n;
        return 0;
    }

    void destroy_lanes()
    {
        for (int j = 1; j < kServerLanes; j++) {
            if (lane_s[j]) (void)hipStreamSynchronize(lane_s[j]);
            if (lane_s[j]) (void)hipStreamDestroy(lane_s[j]);
            if (lane_ev[j]) (void)hipEventDestroy(lane_ev[j]);
            lane_s[j] = nullptr;
            lane_ev[j] = nullptr;
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        fork_ev = nullptr;
        lanes_busy = false;
        lanes = 1;
    }

    int end_forward(SReq &r)
    {
        join();
        if (r.st.b.flush_last_snapshot(stream)) return fail("last snapshot");
        r.st.b.reset_state(stream);
        return 0;
    }

    int finish(SReq &r)
    {
        join();
        const int nx = r.p.nx, nz = r.p.nz;
        std::vector<float> h(r.st.b.nint);
        hipError_t e = hipMemcpyAsync(h.data(), r.st.b.I, sizeof(float) * r.st.b.nint, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail("image download", e);
        r.image.resize((size_t)nx * nz);
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) r.image[(size_t)i * nz + k] = h[(size_t)k * nx + i];
        return 0;
    }

    int sync()
    {
        join();
        const hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail("stream", e);
        return 0;
    }
};

}  // namespace

struct rtm_server {
    HipBackend be;
    rtm_sched::Core<HipBackend> core;
    explicit rtm_server(int max_slots) : core(be, max_slots) { be.max_slots = max_slots; }
};

extern "C" size_t rtm_server_request_footprint(const rtm_params *p)
{
    if (!p || rtm_params_check(p)) return 0;
    // every buffer is rounded to the arena's 256-byte unit; ~40 buffers
    return sim_bytes(*p, 4) + 40 * 256;
}

static void server_free(rtm_server *s)
{
    s->be.join();
    if (s->be.stream) (void)hipStreamSynchronize(s->be.stream);
    for (auto &r : s->core.slots)
        if (r) s->be.release(*r);
    s->core.slots.clear();
    s->be.trim(0);
    s->be.arena.release();
    s->be.destroy_lanes();
    if (s->be.stream) (void)hipStreamDestroy(s->be.stream);
    s->be.stream = nullptr;
}

extern "C" int rtm_server_create(int device, size_t hbm_bytes, int max_slots, int snapshot_mode, rtm_server **out)
{
    if (!out || max_slots < 1 || hbm_bytes < (1u << 20)) { rtm_set_error("rtm_server_create: bad argument"); return -1; }
    *out = nullptr;
    if (snapshot_mode < RTM_SNAP_FP32 || snapshot_mode > RTM_SNAP_I16 || snapshot_mode == RTM_SNAP_BOUNDARY) {
        rtm_set_error("rtm_server_create: snapshot mode %d not available", snapshot_mode);
        return -1;
    }
    HIP_TRY(hipSetDevice(device));
    rtm_server *s = new rtm_server(max_slots);
    s->be.device = device;
    s->be.snapshot_mode = snapshot_mode;
    s->core.group_bytes = kGroupBytes;
    auto fail = [&]() { server_free(s); delete s; return -1; };
    if (hipStreamCreateWithFlags(&s->be.stream, hipStreamNonBlocking) != hipSuccess) {
        rtm_set_error("rtm_server_create: stream creation failed");
        return fail();
    }
    if (s->be.arena.init(hbm_bytes)) return fail();
    *out = s;
    return 0;
}

extern "C" int rtm_server_destroy(rtm_server *s)
{
    if (!s) return 0;
    (void)hipSetDevice(s->be.device);
    server_free(s);
    delete s;
    return 0;
}

extern "C" int rtm_server_set_tiling(rtm_server *s, int rows_per_thread, int graph_block, int interior, int seg_rows)
{
    if (!s || graph_block < 1) { rtm_set_error("rtm_server_set_tiling: bad argument"); return -1; }
    if (check_shape(rows_per_thread, interior) || check_store_shape(s->be.snapshot_mode, interior)) return -1;
    if (interior == rtm::IN_PAIRS) {
        rtm_set_error("the request server steps tile and column shapes (interior kernel 0-3), not pair items");
        return -1;
    }
    std::lock_guard<std::mutex> l(s->core.tick_mu);
    s->be.rz = rows_per_thread;
    s->core.block = graph_block;
    s->be.interior = interior;
    s->be.seg_rows = seg_rows;
    s->core.lists_dirty = true;
    return 0;
}

extern "C" int rtm_server_submit(rtm_server *s, const rtm_params *p, const float *vpad, const float *wavelet,
                                 const float *seis, int shot_x, int shot_z)
{
    if (!s || !p || !vpad || !wavelet || !seis) { rtm_set_error("rtm_server_submit: null argument"); return -1; }
    if (rtm_params_check(p)) return -1;
    if (p->snapshot_mode != s->be.snapshot_mode) {
        rtm_set_error("rtm_server_submit: snapshot mode %d differs from the server's %d", p->snapshot_mode,
                      s->be.snapshot_mode);
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
    if (rtm_server_request_footprint(p) > s->be.arena.size) {   // the arena's size is fixed at creation
        rtm_set_error("rtm_server_submit: request needs %zu bytes, the arena holds %zu", rtm_server_request_footprint(p),
                      s->be.arena.size);
        return -1;
    }
    // the inputs are copied outside every lock: a submission waits for no tick
    auto r = std::make_unique<SReq>();
    r->p = *p;
    r->vpad.assign(vpad, vpad + (size_t)X * Z);
    r->wav.assign(wavelet, wavelet + p->sampling_wavelet);
    r->seis.assign(seis, seis + (size_t)p->number_of_receivers * p->number_of_time_steps);
    r->sx = shot_x;
    r->sz = shot_z;
    return s->core.submit(std::move(r));
}

extern "C" int rtm_server_tick(rtm_server *s)
{
    if (!s) { rtm_set_error("rtm_server_tick: null server"); return -1; }
    HIP_TRY(hipSetDevice(s->be.device));
    return s->core.tick();
}

extern "C" int rtm_server_next_done(rtm_server *s, int *id)
{
    if (!s || !id) { rtm_set_error("rtm_server_next_done: null argument"); return -1; }
    return s->core.next_done(id);
}

extern "C" int rtm_server_take_image(rtm_server *s, int id, float *image, size_t n)
{
    if (!s || !image) { rtm_set_error("rtm_server_take_image: null argument"); return -1; }
    std::string why;
    const int rc = s->core.take(id, image, n, &why);
    if (rc) rtm_set_error("rtm_server_take_image: %s", why.c_str());
    return rc;
}

extern "C" int rtm_server_set_group_bytes(rtm_server *s, size_t bytes)
{
    if (!s) { rtm_set_error("rtm_server_set_group_bytes: null server"); return -1; }
    std::lock_guard<std::mutex> l(s->core.tick_mu);
    s->core.group_bytes = bytes;
    s->core.lists_dirty = true;
    return 0;
}

extern "C" int rtm_server_set_lanes(rtm_server *s, int lanes)
{
    if (!s || lanes < 1 || lanes > kServerLanes) {
        rtm_set_error("rtm_server_set_lanes: launch groups in flight must be 1 to %d", kServerLanes);
        return -1;
    }
    std::lock_guard<std::mutex> l(s->core.tick_mu);
    HIP_TRY(hipSetDevice(s->be.device));
    if (s->be.set_lanes(lanes)) {
        rtm_set_error("%s", s->be.error().c_str());
        return -1;
    }
    return 0;
}

extern "C" int rtm_server_get_lanes(rtm_server *s, int *lanes)
{
    if (!s || !lanes) { rtm_set_error("rtm_server_get_lanes: bad argument"); return -1; }
    *lanes = s->be.lanes;
    return 0;
}

extern "C" int rtm_server_get_group_bytes(rtm_server *s, size_t *bytes, int *groups)
{
    if (!s) { rtm_set_error("rtm_server_get_group_bytes: null server"); return -1; }
    std::lock_guard<std::mutex> l(s->core.tick_mu);
    if (bytes) *bytes = s->core.group_bytes;
    if (groups) {
        int g = 0;
        for (const auto &m : s->core.groups) g += m.empty() ? 0 : 1;
        *groups = g;
    }
    return 0;
}

extern "C" int rtm_server_stats(rtm_server *s, int *active, int *queued, size_t *bytes_in_use, long long *steps)
{
    if (!s) { rtm_set_error("rtm_server_stats: null server"); return -1; }
    std::lock_guard<std::mutex> l(s->core.tick_mu);
    if (active) *active = s->core.active();
    if (queued) *queued = s->core.queued();
    if (bytes_in_use) *bytes_in_use = s->be.arena.in_use;
    if (steps) *steps = s->core.steps_run;
    return 0;
}

extern "C" int rtm_server_failures(rtm_server *s)
{
    if (!s) { rtm_set_error("rtm_server_failures: null server"); return -1; }
    return s->core.failures.load();
}
