// rtm_sched.h -- the request server's bookkeeping, independent of the device:
// the arrival queue, the request slots, launch groups, the pass changes at
// block boundaries, completions and failures.  rtm_serve.hip runs it over
// its HIP backend (device arena, executors, graph launches); the CPU tests
// run it over a stand-in backend whose allocator and launches fail on demand
// (tests/cpu_serve_harness.cpp).
//
// Backend B provides (every int result: 0 success, < 0 error with the
// reason in B::error()):
//   using State                          per-request device state
//   int admit(Req &r, int slot)          allocate, upload the inputs, reset the
//                                        state for the forward pass; 1 = no
//                                        memory now (nothing is held on return)
//   void release(Req &r)                 free the request's memory (idempotent)
//   size_t step_bytes(const Req &r)      per-step working set (launch groups)
//   int prepare(int g, members)          executor g's work lists for `members`
//   void trim(int ngroups)               drop executors ngroups and up
//   void reset(int g)                    drop executor g after its failure
//   int upload(int g, slots)             descriptors and step counters of
//                                        every slot (null: empty) to executor g
//   int run(int g, int steps)            executor g advances its members
//   int end_forward(Req &r)              forward pass over: set up the backward
//   int finish(Req &r)                   the image to r.image ([Nx][Nz])
//   int sync()                           wait for the device's queued work
//   std::string error()                  the last failure's reason
//
// Failures.  A failing step of one launch group fails that group's requests
// only: they are retired with the reason (their memory released once the
// device has drained), the group's executor is dropped, and the other
// groups, the queue and later ticks carry on.  A failed admission (an upload
// error, or a request that does not fit the empty arena) fails that request
// alone.  A sticky device error fails every later launch, so then every
// request fails in turn, each with the reason.
//
// Threads.  submit(), next_done() and take() may be called from any thread
// at any time: they touch only the queue and the done list, under q_mu.
// tick(), stats() and the configuration calls serialise on tick_mu, so a
// submission never waits for a tick's launches.
#ifndef RTM_SCHED_H
#define RTM_SCHED_H

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "rtm_hip.h"

namespace rtm_sched {

enum Pass : int { P_QUEUED = 0, P_FORWARD = 1, P_BACKWARD = 2 };

template <class State>
struct Req {
    int id = -1;
    rtm_params p{};
    std::vector<float> vpad, wav, seis;
    int sx = 0, sz = 0;
    int slot = -1;
    int pass = P_QUEUED;
    int n = 0;                   // steps done in the current pass
    bool failed = false;
    std::string error;           // the failure's reason
    std::vector<float> image;    // [Nx][Nz] once complete
    State st;                    // the backend's per-request state
};

template <class B>
struct Core {
    using R = Req<typename B::State>;

    B &be;
    int max_slots;
    int block = 100;                      // steps per tick
    size_t group_bytes = 0;               // launch-group budget (0: one group)

    std::mutex q_mu;                      // queue, done, next_id
    std::deque<std::unique_ptr<R>> queue;
    std::deque<std::unique_ptr<R>> done;
    int next_id = 0;

    std::mutex tick_mu;                   // everything below
    std::vector<std::unique_ptr<R>> slots;
    std::vector<std::vector<int>> groups; // slot indices per launch group
    bool lists_dirty = true, descs_dirty = true;
    long long steps_run = 0;
    std::atomic<int> failures{0};

    Core(B &backend, int nslots) : be(backend), max_slots(nslots), slots(nslots) {}

    // queue a request (its inputs already copied into it); returns its id
    int submit(std::unique_ptr<R> r)
    {
        std::lock_guard<std::mutex> l(q_mu);
        r->id = next_id++;
        r->pass = P_QUEUED;
        const int id = r->id;
        queue.push_back(std::move(r));
        return id;
    }

    int active() const
    {
        int a = 0;
        for (const auto &r : slots) a += r ? 1 : 0;
        return a;
    }

    // retire request r (already out of its slot) with the reason
    void retire_failed(std::unique_ptr<R> r, const std::string &why)
    {
        be.release(*r);
        r->failed = true;
        r->error = why.empty() ? std::string("device error") : why;
        r->slot = -1;
        r->vpad.clear(); r->vpad.shrink_to_fit();
        r->seis.clear(); r->seis.shrink_to_fit();
        failures++;
        std::lock_guard<std::mutex> l(q_mu);
        done.push_back(std::move(r));
    }

    void fail_slot(int slot, const std::string &why)
    {
        std::unique_ptr<R> r = std::move(slots[slot]);
        if (!r) return;
        retire_failed(std::move(r), why);
        lists_dirty = descs_dirty = true;
    }

    // a launch group failed: its requests fail, the others go on
    void fail_group(int g, const std::string &why)
    {
        (void)be.sync();             // its launches may still run: drain before freeing
        be.reset(g);
        for (int s : groups[g]) fail_slot(s, why);
        groups[g].clear();
    }

    // admit queued requests in arrival order while a slot is free and the
    // backend has room; a request the empty device cannot hold, or whose
    // uploads fail, fails alone
    void admit()
    {
        for (;;) {
            int slot = -1;
            for (int i = 0; i < max_slots && slot < 0; i++)
                if (!slots[i]) slot = i;
            if (slot < 0) return;
            std::unique_ptr<R> r;
            {
                std::lock_guard<std::mutex> l(q_mu);
                if (queue.empty()) return;
                r = std::move(queue.front());
                queue.pop_front();
            }
            const int rc = be.admit(*r, slot);
            if (rc == 1) {
                be.release(*r);
                if (active() == 0) {   // nothing else holds memory: waiting cannot help
                    const std::string why = "request " + std::to_string(r->id) + " does not fit the empty arena";
                    retire_failed(std::move(r), why);
                    continue;
                }
                std::lock_guard<std::mutex> l(q_mu);   // wait for memory to free up
                queue.push_front(std::move(r));
                return;
            }
            if (rc != 0) {
                const std::string why = be.error();
                (void)be.sync();     // uploads in flight read the request's host inputs
                retire_failed(std::move(r), why);
                continue;
            }
            r->slot = slot;
            r->pass = P_FORWARD;
            r->n = 0;
            slots[slot] = std::move(r);
            lists_dirty = descs_dirty = true;
        }
    }

    // launch groups: greedily in slot order, per-step working sets summing
    // to at most group_bytes; each group's executor gets its work lists.
    // With a budget, the forward-pass requests come first and a group never
    // holds both passes: a launch then has one pass's store or imaging steps
    // (one step in TS, not two), and the other steps run the light tile
    // instance.
    void regroup()
    {
        // the last tick's launches may still read the executors' lists
        if (be.sync()) {
            const std::string why = be.error();
            for (int i = 0; i < max_slots; i++) fail_slot(i, why);
            groups.clear();
            be.trim(0);
            lists_dirty = false;
            return;
        }
        groups.clear();
        size_t sum = 0;
        int last_pass = -1;
        for (int pass : {(int)P_FORWARD, (int)P_BACKWARD})
            for (int i = 0; i < max_slots; i++) {
                if (!slots[i] || (group_bytes && slots[i]->pass != pass) || (!group_bytes && pass != P_FORWARD))
                    continue;
                const size_t w = be.step_bytes(*slots[i]);
                if (groups.empty() || (group_bytes && (sum + w > group_bytes || slots[i]->pass != last_pass))) {
                    groups.emplace_back();
                    sum = 0;
                }
                groups.back().push_back(i);
                sum += w;
                last_pass = slots[i]->pass;
            }
        be.trim((int)groups.size());
        for (int g = 0; g < (int)groups.size(); g++) {
            std::vector<const R *> m;
            for (int s : groups[g]) m.push_back(slots[s].get());
            if (be.prepare(g, m)) fail_group(g, be.error());
        }
        lists_dirty = false;
        descs_dirty = true;
    }

    std::vector<const R *> slot_view() const
    {
        std::vector<const R *> v(max_slots, nullptr);
        for (int i = 0; i < max_slots; i++) v[i] = slots[i].get();
        return v;
    }

    // one tick: admit, advance every active request one block, then the
    // pass changes and completions; returns the requests queued or active
    int tick()
    {
        std::lock_guard<std::mutex> tl(tick_mu);
        admit();
        if (active()) {
            if (lists_dirty) regroup();
            if (descs_dirty) {
                const auto view = slot_view();
                for (int g = 0; g < (int)groups.size(); g++)
                    if (!groups[g].empty() && be.upload(g, view)) fail_group(g, be.error());
                descs_dirty = false;
            }
            bool ran = false;
            for (int g = 0; g < (int)groups.size(); g++) {
                if (groups[g].empty()) continue;
                if (be.run(g, block)) fail_group(g, be.error());
                else ran = true;
            }
            if (ran) steps_run += block;
            for (int i = 0; i < max_slots; i++) {
                R *r = slots[i].get();
                if (!r) continue;
                r->n += block;
                if (r->n < r->p.number_of_time_steps) continue;
                if (r->pass == P_FORWARD) {
                    if (be.end_forward(*r)) {
                        const std::string why = be.error();
                        (void)be.sync();
                        fail_slot(i, why);
                        continue;
                    }
                    r->pass = P_BACKWARD;
                    r->n = 0;
                    descs_dirty = true;
                    if (group_bytes) lists_dirty = true;   // groups hold one pass each
                } else {
                    const int rc = be.finish(*r);
                    std::unique_ptr<R> fin = std::move(slots[i]);
                    lists_dirty = descs_dirty = true;
                    if (rc) {
                        const std::string why = be.error();
                        (void)be.sync();
                        retire_failed(std::move(fin), why);
                        continue;
                    }
                    be.release(*fin);
                    fin->slot = -1;
                    fin->vpad.clear(); fin->vpad.shrink_to_fit();
                    fin->seis.clear(); fin->seis.shrink_to_fit();
                    std::lock_guard<std::mutex> l(q_mu);
                    done.push_back(std::move(fin));
                }
            }
        }
        std::lock_guard<std::mutex> l(q_mu);
        return (int)queue.size() + active();
    }

    // 1 and *id = the oldest retired request not yet taken, 0 if none
    int next_done(int *id)
    {
        std::lock_guard<std::mutex> l(q_mu);
        if (done.empty()) return 0;
        *id = done.front()->id;
        return 1;
    }

    // a retired request's image into `image` (n = Nx*Nz), and forget it:
    // 0, or -2 for a failed request (its reason in *why), -1 if unknown or
    // the size differs (reason in *why)
    int take(int id, float *image, size_t n, std::string *why)
    {
        std::lock_guard<std::mutex> l(q_mu);
        for (auto it = done.begin(); it != done.end(); ++it) {
            if ((*it)->id != id) continue;
            if ((*it)->failed) {
                *why = "request " + std::to_string(id) + " failed: " + (*it)->error;
                done.erase(it);
                return -2;
            }
            if (n != (*it)->image.size()) {
                *why = "request " + std::to_string(id) + " has " + std::to_string((*it)->image.size()) +
                       " image points, not " + std::to_string(n);
                return -1;
            }
            std::copy((*it)->image.begin(), (*it)->image.end(), image);
            done.erase(it);
            return 0;
        }
        *why = "request " + std::to_string(id) + " is not complete";
        return -1;
    }

    int queued()
    {
        std::lock_guard<std::mutex> l(q_mu);
        return (int)queue.size();
    }
};

}  // namespace rtm_sched

#endif /* RTM_SCHED_H */
