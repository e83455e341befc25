// Host-only check of the serving loop's scheduler (blitzen_amd/csrc/server_core.h) over a fake device, built with
// ThreadSanitizer and, separately, AddressSanitizer + UBSan. Several client threads submit, cancel, wait, fetch and
// release while the launcher and completer threads batch and complete; the fake device finishes each batch after a
// random delay and can fail a chosen launch or poll, or report a given-up wait. Checked:
//   - every submitted request ends exactly once, with OK (and the count the device produced for it), CANCELLED,
//     TIMEOUT (deadline already past), or the error of the launch / poll that failed its batch (BZ_E_INTERNAL for a
//     given-up wait), and nothing else;
//   - requests are admitted in submission order; a batch holds at most max_batch_requests requests and at most
//     max_inflight_batches batches are on the device at once;
//   - the working and output regions of live requests never overlap, never exceed the budget, and the budget is
//     fully returned once everything is fetched or released;
//   - stop() cancels what was never admitted and returns with nothing left on the device.
// Prints "ok" or the first failure.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "../../blitzen_amd/csrc/server_core.h"

using bz::Arena;

static std::atomic<int> g_failures{ 0 };
static std::mutex g_print_mu;

#define EXPECT(cond, ...)                                                    \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::lock_guard<std::mutex> g_(g_print_mu);                      \
            if (g_failures.fetch_add(1) == 0) {                              \
                std::printf("FAIL %s:%d %s: ", __FILE__, __LINE__, #cond);   \
                std::printf(__VA_ARGS__);                                    \
                std::printf("\n");                                           \
            }                                                                \
        }                                                                    \
    } while (0)

struct FakeDevice;
using Req = bz::SchedReq<FakeDevice>;
using ReqP = std::shared_ptr<Req>;

struct FakeDevice {
    struct ReqData {
        uint32_t expect = 0;  // the count this request's cull produces
    };
    struct SlotState {
        std::vector<ReqP> batch;
        uint64_t due = 0;
        bool active = false;
    };
    std::mutex mu;
    std::vector<SlotState> slots;
    std::map<uint64_t, uint64_t> live;  // arena offset -> bytes of every region a launched request holds
    uint32_t max_batch = 0;
    uint64_t budget = 0;
    uint64_t launches = 0, polls = 0, last_ticket_admitted = 0;
    uint64_t fail_launch = ~0ull, fail_poll = ~0ull, give_up_launch = ~0ull;
    std::set<uint32_t> give_up_slots;
    std::mt19937_64 rng{ 7 };
    uint32_t launch_delay_us = 0;  // time a launch takes (widens the window between admission and flight)

    FakeDevice(uint32_t n_slots, uint32_t max_batch_, uint64_t budget_) : slots(n_slots), max_batch(max_batch_),
                                                                          budget(budget_) {}
    void thread_init() {}

    void add_region(uint64_t off, uint64_t n)
    {
        n = Arena::round_up(n ? n : 1, 256);
        EXPECT(off + n <= budget, "region [%llu, +%llu) beyond the budget", (unsigned long long)off,
               (unsigned long long)n);
        auto it = live.lower_bound(off);
        if (it != live.end()) EXPECT(off + n <= it->first, "region at %llu overlaps", (unsigned long long)off);
        if (it != live.begin()) {
            auto p = std::prev(it);
            EXPECT(p->first + p->second <= off, "region at %llu overlaps", (unsigned long long)off);
        }
        live[off] = n;
    }

    bz_status launch(uint32_t slot, const std::vector<ReqP>& batch)
    {
        if (launch_delay_us) std::this_thread::sleep_for(std::chrono::microseconds(launch_delay_us));
        std::lock_guard<std::mutex> g(mu);
        ++launches;
        EXPECT(!batch.empty() && batch.size() <= max_batch, "batch of %zu", batch.size());
        EXPECT(!slots[slot].active, "slot %u launched while in flight", slot);
        for (const ReqP& q : batch) {
            EXPECT(q->state == bz::RQ_INFLIGHT, "request %llu launched in state %d", (unsigned long long)q->ticket,
                   q->state);
            EXPECT(q->ticket > last_ticket_admitted, "request %llu admitted after %llu", (unsigned long long)q->ticket,
                   (unsigned long long)last_ticket_admitted);
            last_ticket_admitted = q->ticket;
            add_region(q->work_off, q->work_bytes);
            add_region(q->out_off, q->out_bytes);
        }
        if (launches == fail_launch) {
            for (const ReqP& q : batch) live.erase(q->work_off), live.erase(q->out_off);  // never ran
            return BZ_E_HIP;
        }
        slots[slot].batch = batch;
        slots[slot].due = bz::now_ns() + rng() % 300000;
        slots[slot].active = true;
        if (launches == give_up_launch) give_up_slots.insert(slot);
        return BZ_OK;
    }

    bz_status poll(uint32_t slot, bool* done)
    {
        std::lock_guard<std::mutex> g(mu);
        EXPECT(slots[slot].active, "poll of idle slot %u", slot);
        *done = false;
        if (++polls == fail_poll) {
            release_work(slot);
            for (const ReqP& q : slots[slot].batch) live.erase(q->out_off);  // released at completion
            slots[slot].active = false;
            return BZ_E_HIP;
        }
        if (bz::now_ns() >= slots[slot].due) *done = true;
        return BZ_OK;
    }

    void release_work(uint32_t slot)
    {
        for (const ReqP& q : slots[slot].batch) live.erase(q->work_off);
    }

    void collect(uint32_t slot, uint32_t n, uint32_t* counts, bool* gave_up)
    {
        std::lock_guard<std::mutex> g(mu);
        EXPECT(n == slots[slot].batch.size(), "collect of %u for a batch of %zu", n, slots[slot].batch.size());
        for (uint32_t k = 0; k < n; ++k) counts[k] = slots[slot].batch[k]->dev.expect;
        *gave_up = give_up_slots.erase(slot) > 0;
        release_work(slot);
        for (const ReqP& q : slots[slot].batch)
            if (q->cancel) live.erase(q->out_off);  // a cancelled request's outputs are released at completion
        slots[slot].active = false;
        slots[slot].batch.clear();
    }

    void released(uint64_t out_off)
    {
        std::lock_guard<std::mutex> g(mu);
        live.erase(out_off);
    }
};

struct Outcome {
    bz_status status = 1;
    uint32_t count = 0;
    int seen = 0;
    bool cancelled = false, past_deadline = false;
    uint32_t expect = 0;
};

// One scenario: `clients` threads each submit `per_client` requests.
static void scenario(const char* name, uint32_t inflight, uint32_t max_batch, uint64_t budget, int clients,
                     int per_client, uint64_t fail_launch, uint64_t fail_poll, uint64_t give_up_launch,
                     bz_status expect_fail)
{
    bz_server_cfg cfg{};
    cfg.hbm_budget_bytes = budget;
    cfg.max_inflight_batches = inflight;
    cfg.max_batch_requests = max_batch;
    auto dev_p = std::make_unique<FakeDevice>(inflight, max_batch, budget);
    FakeDevice& dev = *dev_p;
    dev.fail_launch = fail_launch;
    dev.fail_poll = fail_poll;
    dev.give_up_launch = give_up_launch;
    // mutexes on the heap: a fresh address per scenario (ThreadSanitizer does not see a std::mutex destroyed, so a
    // stack address reused by the next scenario would look like the same mutex)
    auto sched_p = std::make_unique<bz::Scheduler<FakeDevice>>(dev, cfg, 1u << 20, 256);
    bz::Scheduler<FakeDevice>& sched = *sched_p;
    sched.start();
    auto out_mu_p = std::make_unique<std::mutex>();
    std::mutex& out_mu = *out_mu_p;
    std::map<uint64_t, Outcome> out;
    std::atomic<uint64_t> submitted{ 0 };
    auto client = [&](int id) {
        std::mt19937_64 rng(1000 + id);
        std::vector<uint64_t> mine;
        for (int i = 0; i < per_client; ++i) {
            auto q = std::make_shared<Req>();
            q->n_objects = 1 + (uint32_t)(rng() % 200000);
            q->work_bytes = 4096 + rng() % (budget / 8);
            q->out_bytes = 256 + rng() % (budget / 16);
            q->dev.expect = (uint32_t)(rng() % 100000);
            const bool past = rng() % 13 == 0;
            q->r.deadline_ns = past ? 1 : 0;
            q->r.user_tag = ((uint64_t)id << 32) | (uint64_t)i;
            uint64_t t = 0;
            const bz_status s = sched.submit(q, 0, &t);
            EXPECT(s == BZ_OK, "submit %s", name);
            if (s) continue;
            ++submitted;
            bool cancelled = false;
            if (rng() % 11 == 0) cancelled = sched.cancel(t) == BZ_OK;
            {
                std::lock_guard<std::mutex> g(out_mu);
                Outcome& o = out[t];
                o.cancelled = cancelled;
                o.past_deadline = past;
                o.expect = q->dev.expect;
            }
            mine.push_back(t);
            if (rng() % 3 == 0 && !mine.empty()) {  // resolve an earlier one now and then
                const uint64_t w = mine.front();
                mine.erase(mine.begin());
                bz_result r{};
                EXPECT(sched.wait(w, 0, &r) == BZ_OK, "wait");
                ReqP taken;
                EXPECT(sched.take(w, taken) == BZ_OK, "take");
                if (taken) {
                    dev.released(taken->out_off);
                    sched.release(*taken);
                }
                std::lock_guard<std::mutex> g(out_mu);
                out[w].status = r.status;
                out[w].count = r.count;
                out[w].seen++;
            }
        }
        for (uint64_t w : mine) {
            bz_result r{};
            EXPECT(sched.wait(w, 0, &r) == BZ_OK, "wait");
            EXPECT(sched.poll(w, &r) == BZ_OK, "poll after wait");
            ReqP taken;
            EXPECT(sched.take(w, taken) == BZ_OK, "take");
            EXPECT(sched.take(w, taken) == BZ_E_NOT_FOUND, "second take");
            if (taken) {
                dev.released(taken->out_off);
                sched.release(*taken);
            }
            std::lock_guard<std::mutex> g(out_mu);
            out[w].status = r.status;
            out[w].count = r.count;
            out[w].seen++;
        }
    };
    std::vector<std::thread> th;
    for (int c = 0; c < clients; ++c) th.emplace_back(client, c);
    for (auto& t : th) t.join();
    const bz_server_stats st = sched.stats();
    int n_fail = 0;
    for (auto& kv : out) {
        const Outcome& o = kv.second;
        EXPECT(o.seen == 1, "%s: ticket %llu seen %d times", name, (unsigned long long)kv.first, o.seen);
        if (o.status == BZ_OK) {
            EXPECT(o.count == o.expect, "%s: count %u expected %u", name, o.count, o.expect);
            EXPECT(!o.cancelled && !o.past_deadline, "%s: a cancelled or late request succeeded", name);
        } else if (o.status == BZ_E_CANCELLED) {
            EXPECT(o.cancelled, "%s: cancelled without cancel", name);
        } else if (o.status == BZ_E_TIMEOUT) {
            EXPECT(o.past_deadline, "%s: timeout without deadline", name);
        } else {
            EXPECT(o.status == expect_fail, "%s: status %d", name, o.status);
            ++n_fail;
        }
    }
    EXPECT((expect_fail == BZ_OK) == (n_fail == 0), "%s: %d failed requests", name, n_fail);
    EXPECT(st.submitted == submitted.load(), "%s: submitted %llu", name, (unsigned long long)st.submitted);
    EXPECT(st.budget_in_use == 0, "%s: %llu bytes still in use", name, (unsigned long long)st.budget_in_use);
    EXPECT(st.budget_peak <= budget && st.budget_peak > 0, "%s: peak %llu", name,
           (unsigned long long)st.budget_peak);
    EXPECT(st.inflight_batches == 0 && st.queue_depth == 0, "%s: left busy", name);
    {
        std::lock_guard<std::mutex> g(dev.mu);
        EXPECT(dev.live.empty(), "%s: %zu regions left", name, dev.live.size());
    }
    sched.stop();
}

// stop() with requests still queued behind a batch that holds the whole budget: they end CANCELLED, and a batch
// being launched while stop() runs still completes.
static void stop_with_queue(uint32_t launch_delay_us)
{
    bz_server_cfg cfg{};
    cfg.hbm_budget_bytes = 1 << 20;
    cfg.max_inflight_batches = 1;
    cfg.max_batch_requests = 4;
    auto dev_p = std::make_unique<FakeDevice>(1, 4, cfg.hbm_budget_bytes);
    FakeDevice& dev = *dev_p;
    dev.launch_delay_us = launch_delay_us;
    std::vector<ReqP> qs;
    std::vector<uint64_t> tickets;
    {
        auto sched_p = std::make_unique<bz::Scheduler<FakeDevice>>(dev, cfg, 1u << 20, 256);
        bz::Scheduler<FakeDevice>& sched = *sched_p;
        sched.start();
        for (int i = 0; i < 20; ++i) {
            auto q = std::make_shared<Req>();
            q->n_objects = 10;
            q->work_bytes = 600 << 10;  // one request at a time fits
            q->out_bytes = 1024;
            uint64_t t = 0;
            EXPECT(sched.submit(q, 0, &t) == BZ_OK, "submit");
            qs.push_back(q);
            tickets.push_back(t);
        }
        sched.stop();
        bz_result r{};
        int cancelled = 0, ok = 0;
        for (uint64_t t : tickets) {
            EXPECT(sched.poll(t, &r) == BZ_OK, "poll after stop");
            cancelled += r.status == BZ_E_CANCELLED;
            ok += r.status == BZ_OK;
        }
        EXPECT(cancelled + ok == 20 && cancelled > 0, "stop: %d cancelled, %d ok", cancelled, ok);
    }
    std::lock_guard<std::mutex> g(dev.mu);
    for (auto& s : dev.slots) EXPECT(!s.active, "a batch left on the device after stop");
}

int main()
{
    // a hang fails the check instead of stalling it (untimed waits: this ThreadSanitizer does not intercept the
    // clock-based condition-variable wait that a timed wait uses)
    std::atomic<bool> finished{ false };
    std::thread watchdog([&] {
        for (int i = 0; i < 3000 && !finished.load(); ++i) std::this_thread::sleep_for(std::chrono::milliseconds(100));
        if (!finished.load()) {
            std::printf("FAIL hang: no progress for 300 s\n");
            std::fflush(stdout);
            std::_Exit(3);
        }
    });
    scenario("plain", 4, 16, 64ull << 20, 6, 300, ~0ull, ~0ull, ~0ull, BZ_OK);
    scenario("tight budget", 3, 8, 4ull << 20, 4, 200, ~0ull, ~0ull, ~0ull, BZ_OK);
    scenario("launch fails", 4, 16, 64ull << 20, 4, 200, 3, ~0ull, ~0ull, BZ_E_HIP);
    scenario("poll fails", 2, 4, 16ull << 20, 3, 150, ~0ull, 5, ~0ull, BZ_E_HIP);
    scenario("wait given up", 4, 16, 64ull << 20, 4, 200, ~0ull, ~0ull, 2, BZ_E_INTERNAL);
    scenario("one slot, batches of one", 1, 1, 8ull << 20, 3, 100, ~0ull, ~0ull, ~0ull, BZ_OK);
    for (int i = 0; i < 20; ++i) stop_with_queue(i % 2 ? 300 : 0);
    finished = true;
    watchdog.join();
    if (g_failures.load() == 0) std::printf("ok\n");
    return g_failures.load() ? 1 : 0;
}
