// cpu_serve_harness.cpp -- the request server's scheduling core
// (rtmuq2d_amd/csrc/rtm_sched.h) over a stand-in backend, on the CPU.
//
// The stand-in allocator counts bytes against a capacity and can refuse a
// request (no room) or fail its uploads; the stand-in launches advance each
// member's step count and can fail a chosen launch group at a chosen tick,
// fail its work-list build, fail a pass switch or an image download, or fail
// everything from some tick on (a sticky device error).  A request's "image"
// is its id and the steps its group ran for it, so a result shows that the
// request ran exactly its own steps, whatever failed around it.
//
// The stand-in also checks the ordering rules the device needs: no memory is
// released, and no work list rebuilt, while a launch or upload that may read
// it is still in flight (only sync() retires in-flight work).
//
// usage: cpu_serve_harness SCENARIO   -- prints "ok" and exits 0, or the
// first violated expectation and exits 1.  Test infrastructure only
// (tests/test_serve_core.py builds and runs it).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "rtm_sched.h"

namespace {

struct FakeState {
    size_t bytes = 0;
    bool held = false;
};
using FReq = rtm_sched::Req<FakeState>;

int g_fail = 0;
#define EXPECT(c)                                                                  \
    do {                                                                           \
        if (!(c)) {                                                                \
            std::fprintf(stderr, "expectation failed: %s (line %d)\n", #c, __LINE__); \
            g_fail = 1;                                                            \
        }                                                                          \
    } while (0)

struct Fake {
    using State = FakeState;
    size_t cap = 1 << 30, in_use = 0, peak = 0;
    int tick = 0;                          // set by the driver before each tick
    std::set<int> refuse_ids, fail_upload_ids, fail_end_forward_ids, fail_finish_ids;
    int fail_run_tick = -1, fail_run_group = -1;
    int fail_prepare_tick = -1, fail_prepare_group = -1;
    int sticky_from = -1;                  // every run and upload fails from this tick on
    std::map<int, long long> steps;        // request id -> steps run for it
    std::vector<std::vector<int>> group_ids;
    std::set<int> in_flight;               // request ids whose device work may still run
    bool lists_in_flight = false;
    int violations = 0;
    int mixed_pass_groups = 0;             // prepared groups holding both passes
    std::set<int> passes_grouped;          // passes seen in prepared groups
    std::string err;

    std::string error() const { return err; }
    static size_t footprint(const rtm_params &p)
    {
        return (size_t)16 * (p.nx + p.thickness_left + p.thickness_right) *
               (p.nz + p.thickness_up + p.thickness_down);
    }
    int admit(FReq &r, int)
    {
        if (refuse_ids.count(r.id)) return 1;
        const size_t b = footprint(r.p);
        if (in_use + b > cap) return 1;
        in_use += b;
        peak = in_use > peak ? in_use : peak;
        r.st.bytes = b;
        r.st.held = true;
        in_flight.insert(r.id);            // its uploads are queued
        if (fail_upload_ids.count(r.id) || (sticky_from >= 0 && tick >= sticky_from)) {
            err = "velocity upload: injected";
            return -1;
        }
        return 0;
    }
    void release(FReq &r)
    {
        if (!r.st.held) return;
        if (in_flight.count(r.id)) violations++;
        in_use -= r.st.bytes;
        r.st.held = false;
    }
    size_t step_bytes(const FReq &r) const { return footprint(r.p); }
    int prepare(int g, const std::vector<const FReq *> &m)
    {
        if (lists_in_flight) violations++;
        for (const FReq *r : m) {
            passes_grouped.insert(r->pass);
            if (r->pass != m.front()->pass) {
                mixed_pass_groups++;
                break;
            }
        }
        if ((int)group_ids.size() <= g) group_ids.resize(g + 1);
        group_ids[g].clear();
        for (const FReq *r : m) group_ids[g].push_back(r->id);
        if (tick == fail_prepare_tick && g == fail_prepare_group) {
            err = "work lists: injected";
            return -1;
        }
        return 0;
    }
    void trim(int n)
    {
        if (lists_in_flight) violations++;
        if ((int)group_ids.size() > n) group_ids.resize(n);
    }
    void reset(int g)
    {
        if (g < (int)group_ids.size()) group_ids[g].clear();
    }
    int upload(int g, const std::vector<const FReq *> &slots)
    {
        (void)g;
        (void)slots;
        if (sticky_from >= 0 && tick >= sticky_from) {
            err = "descriptors: injected sticky error";
            return -1;
        }
        return 0;
    }
    int run(int g, int n)
    {
        lists_in_flight = true;
        for (int id : group_ids[g]) in_flight.insert(id);
        if ((tick == fail_run_tick && g == fail_run_group) || (sticky_from >= 0 && tick >= sticky_from)) {
            err = "launch group: injected";
            return -1;
        }
        for (int id : group_ids[g]) steps[id] += n;
        return 0;
    }
    int end_forward(FReq &r)
    {
        if (fail_end_forward_ids.count(r.id)) {
            err = "last snapshot: injected";
            return -1;
        }
        return 0;
    }
    int finish(FReq &r)
    {
        if (fail_finish_ids.count(r.id)) {
            err = "image download: injected";
            return -1;
        }
        r.image.assign((size_t)r.p.nx * r.p.nz, float(r.id * 100000 + steps[r.id]));
        in_flight.clear();                 // the download waits for the stream
        lists_in_flight = false;
        return 0;
    }
    int sync()
    {
        in_flight.clear();
        lists_in_flight = false;
        return (sticky_from >= 0 && tick >= sticky_from) ? (err = "stream: injected sticky error", -1) : 0;
    }
};

using Core = rtm_sched::Core<Fake>;

rtm_params params(int nx, int nz, int nt)
{
    rtm_params p;
    std::memset(&p, 0, sizeof p);
    p.nx = nx; p.nz = nz;
    p.thickness_left = p.thickness_right = p.thickness_up = p.thickness_down = 10;
    p.number_of_time_steps = nt;
    return p;
}

int submit(Core &c, int nx, int nz, int nt)
{
    auto r = std::make_unique<FReq>();
    r->p = params(nx, nz, nt);
    r->vpad.assign(16, 1.f);
    r->seis.assign(16, 1.f);
    return c.submit(std::move(r));
}

// steps a request runs in each pass: whole blocks
long long pass_steps(int nt, int block) { return (long long)(nt + block - 1) / block * block; }

struct Outcome {
    std::map<int, float> image;            // id -> image value (completed)
    std::map<int, std::string> failed;     // id -> reason
};

// tick until nothing is queued or active; collects every retired request
Outcome drain(Core &c, Fake &f, int max_ticks = 10000, std::map<int, std::vector<int>> *late = nullptr)
{
    Outcome o;
    for (int t = 0; t < max_ticks; t++) {
        f.tick++;
        if (late && late->count(f.tick))
            for (int nx : (*late)[f.tick]) submit(c, nx, 20, 300);
        const int left = c.tick();
        int id;
        while (c.next_done(&id) == 1) {
            std::vector<float> img;
            std::string why;
            // size of this request's image is unknown here: ask with 0 first
            const int rc0 = c.take(id, nullptr, 0, &why);
            if (rc0 == -2) {
                o.failed[id] = why;
                continue;
            }
            // the size message names the image's points
            size_t n = 0;
            const char *q = std::strstr(why.c_str(), " has ");
            if (q) n = std::strtoull(q + 5, nullptr, 10);
            img.resize(n);
            EXPECT(c.take(id, img.data(), n, &why) == 0);
            o.image[id] = img.empty() ? -1.f : img[0];
        }
        if (left == 0 && (!late || late->upper_bound(f.tick) == late->end())) break;
    }
    return o;
}

bool done_ok(const Outcome &o, int id, int nt, int block)
{
    auto it = o.image.find(id);
    return it != o.image.end() && it->second == float(id * 100000 + 2 * pass_steps(nt, block));
}

void end_checks(const Core &c, const Fake &f)
{
    EXPECT(f.violations == 0);
    EXPECT(f.in_use == 0);
    EXPECT(c.active() == 0);
}

// ---------------------------------------------------------------- scenarios

// with a group budget, a launch group holds requests of one pass only (the
// groups are rebuilt when a request switches to its backward pass); with
// budget 0 every active request shares one group
void s_pass_groups()
{
    for (size_t budget : {(size_t)1 << 40, (size_t)0}) {
        Fake f;
        Core c(f, 4);
        c.block = 100;
        c.group_bytes = budget;
        std::vector<int> ids, nts;
        for (int k = 0; k < 8; k++) {
            nts.push_back(200 + 100 * k);
            ids.push_back(submit(c, 30, 20, nts.back()));
        }
        const Outcome o = drain(c, f);
        EXPECT(o.failed.empty());
        for (int k = 0; k < 8; k++) EXPECT(done_ok(o, ids[k], nts[k], 100));
        EXPECT(f.passes_grouped.size() == 2);
        if (budget) EXPECT(f.mixed_pass_groups == 0);
        else EXPECT(f.mixed_pass_groups > 0);
        end_checks(c, f);
    }
}

void s_basic()
{
    Fake f;
    Core c(f, 4);
    c.block = 100;
    std::vector<int> ids, nts;
    for (int k = 0; k < 10; k++) {
        nts.push_back(200 + 50 * k);
        ids.push_back(submit(c, 30 + 5 * k, 20, nts.back()));
    }
    const Outcome o = drain(c, f);
    EXPECT(o.failed.empty());
    for (int k = 0; k < 10; k++) EXPECT(done_ok(o, ids[k], nts[k], 100));
    end_checks(c, f);
}

void s_group_failure()
{
    // four requests per group budget: groups {0,1} {2,3} ...; group 1 fails at tick 3
    Fake f;
    Core c(f, 6);
    c.block = 50;
    const size_t one = Fake::footprint(params(40, 20, 1));
    c.group_bytes = 2 * one;
    std::vector<int> ids;
    for (int k = 0; k < 6; k++) ids.push_back(submit(c, 40, 20, 400));
    f.fail_run_tick = 3;
    f.fail_run_group = 1;
    std::map<int, std::vector<int>> late = {{6, {40, 40}}};   // arrive after the failure
    const Outcome o = drain(c, f, 10000, &late);
    EXPECT(o.failed.size() == 2);
    EXPECT(o.failed.count(ids[2]) && o.failed.count(ids[3]));
    for (auto &kv : o.failed) EXPECT(kv.second.find("launch group: injected") != std::string::npos);
    for (int k : {0, 1, 4, 5}) EXPECT(done_ok(o, ids[k], 400, 50));
    EXPECT(done_ok(o, 6, 300, 50) && done_ok(o, 7, 300, 50));   // the server stays usable
    EXPECT(c.failures.load() == 2);
    end_checks(c, f);
}

void s_upload_failure()
{
    Fake f;
    Core c(f, 3);
    c.block = 100;
    std::vector<int> ids;
    for (int k = 0; k < 5; k++) ids.push_back(submit(c, 30, 20, 300));
    f.fail_upload_ids = {ids[1]};
    const Outcome o = drain(c, f);
    EXPECT(o.failed.size() == 1 && o.failed.count(ids[1]));
    EXPECT(o.failed.begin()->second.find("velocity upload") != std::string::npos);
    for (int k : {0, 2, 3, 4}) EXPECT(done_ok(o, ids[k], 300, 100));
    end_checks(c, f);
}

void s_no_fit()
{
    // capacity for two requests: the third waits for memory, then runs; one
    // the empty arena refuses fails alone
    Fake f;
    const size_t one = Fake::footprint(params(30, 20, 1));
    f.cap = 2 * one;
    Core c(f, 8);
    c.block = 100;
    std::vector<int> ids;
    for (int k = 0; k < 3; k++) ids.push_back(submit(c, 30, 20, 300));
    const int big = submit(c, 30, 20, 300);
    f.refuse_ids = {big};
    const Outcome o = drain(c, f);
    EXPECT(o.failed.size() == 1 && o.failed.count(big));
    EXPECT(o.failed.count(big) && o.failed.at(big).find("does not fit the empty arena") != std::string::npos);
    for (int id : ids) EXPECT(done_ok(o, id, 300, 100));
    EXPECT(f.peak <= f.cap);
    end_checks(c, f);
}

void s_prepare_failure()
{
    Fake f;
    Core c(f, 4);
    c.block = 100;
    const size_t one = Fake::footprint(params(30, 20, 1));
    c.group_bytes = 2 * one;
    std::vector<int> ids;
    for (int k = 0; k < 4; k++) ids.push_back(submit(c, 30, 20, 300));
    f.fail_prepare_tick = 1;
    f.fail_prepare_group = 0;
    const Outcome o = drain(c, f);
    EXPECT(o.failed.size() == 2 && o.failed.count(ids[0]) && o.failed.count(ids[1]));
    for (int k : {2, 3}) EXPECT(done_ok(o, ids[k], 300, 100));
    end_checks(c, f);
}

void s_pass_failures()
{
    Fake f;
    Core c(f, 4);
    c.block = 100;
    std::vector<int> ids;
    for (int k = 0; k < 4; k++) ids.push_back(submit(c, 30, 20, 300));
    f.fail_end_forward_ids = {ids[0]};
    f.fail_finish_ids = {ids[3]};
    const Outcome o = drain(c, f);
    EXPECT(o.failed.size() == 2 && o.failed.count(ids[0]) && o.failed.count(ids[3]));
    for (int k : {1, 2}) EXPECT(done_ok(o, ids[k], 300, 100));
    end_checks(c, f);
}

void s_sticky()
{
    // from tick 2 every launch fails: every request fails with a reason,
    // including those submitted later, and tick keeps returning normally
    Fake f;
    Core c(f, 4);
    c.block = 100;
    for (int k = 0; k < 6; k++) submit(c, 30, 20, 300);
    f.sticky_from = 2;
    std::map<int, std::vector<int>> late = {{5, {30}}};
    const Outcome o = drain(c, f, 100, &late);
    EXPECT(o.image.empty());
    EXPECT(o.failed.size() == 7);
    EXPECT(f.violations == 0);
    EXPECT(f.in_use == 0);
}

void s_threads()
{
    // submitters on four threads while one thread ticks and collects
    Fake f;
    Core c(f, 8);
    c.block = 100;
    const int per = 40, nthreads = 4;
    std::vector<std::thread> th;
    std::vector<std::vector<int>> got(nthreads);
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([&, t]() {
            for (int k = 0; k < per; k++) got[t].push_back(submit(c, 20 + (k % 7), 20, 200 + 100 * (k % 3)));
        });
    Outcome o;
    int collected = 0;
    while (collected < per * nthreads) {
        f.tick++;
        c.tick();
        int id;
        while (c.next_done(&id) == 1) {
            std::string why;
            c.take(id, nullptr, 0, &why);
            size_t n = 0;
            const char *q = std::strstr(why.c_str(), " has ");
            if (q) n = std::strtoull(q + 5, nullptr, 10);
            std::vector<float> img(n);
            EXPECT(c.take(id, img.data(), n, &why) == 0);
            o.image[id] = img[0];
            collected++;
        }
    }
    for (auto &t : th) t.join();
    std::set<int> all;
    for (auto &v : got) all.insert(v.begin(), v.end());
    EXPECT((int)all.size() == per * nthreads);
    EXPECT((int)o.image.size() == per * nthreads);
    for (auto &v : got)
        for (size_t k = 0; k < v.size(); k++) EXPECT(done_ok(o, v[k], 200 + 100 * (int)(k % 3), 100));
    end_checks(c, f);
}

}  // namespace

int main(int argc, char **argv)
{
    const std::map<std::string, void (*)()> sc = {
        {"basic", s_basic}, {"group_failure", s_group_failure}, {"upload_failure", s_upload_failure},
        {"no_fit", s_no_fit}, {"prepare_failure", s_prepare_failure}, {"pass_failures", s_pass_failures},
        {"sticky", s_sticky}, {"threads", s_threads}, {"pass_groups", s_pass_groups}};
    if (argc < 2 || !sc.count(argv[1])) {
        std::fprintf(stderr, "usage: %s SCENARIO\n", argv[0]);
        return 2;
    }
    sc.at(argv[1])();
    if (g_fail) return 1;
    std::printf("ok\n");
    return 0;
}
