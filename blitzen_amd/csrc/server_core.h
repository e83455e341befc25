// server_core.h — the serving loop's scheduler, independent of the device: request queue, HBM-budget admission,
// continuous batching over a fixed number of in-flight slots, completion, cancellation, deadlines and statistics.
//
// Threads:
//   clients   submit (any thread; a mutex-protected queue) and poll / wait / take (fetch, release) / cancel.
//   launcher  takes queued requests in arrival order while the budget admits them (one arena region for the
//             request's working memory, one for its outputs), forms a batch of up to max_batch_requests and hands
//             it to the device on a free slot. Up to max_inflight_batches batches run at once, so requests arriving
//             while others are on the device are admitted as soon as a slot and memory are free.
//   completer polls the in-flight slots, stamps completion times, frees the working regions and wakes waiters.
//             Output regions stay allocated until the client fetches or releases them.
//
// The device is a template parameter (no HIP types here), so the scheduling logic is the same code in the library
// (the HIP device in server.cpp) and in the host-only tests (a fake device that can fail a given launch or poll).
// Dev provides:
//   struct ReqData;                                     per-request device description
//   void thread_init();                                 called first on the launcher and completer threads
//   bz_status launch(uint32_t slot, const std::vector<std::shared_ptr<SchedReq<Dev>>>& batch);
//   bz_status poll(uint32_t slot, bool* done);          non-blocking; an error fails the slot's batch
//   void collect(uint32_t slot, uint32_t n, uint32_t* counts, bool* gave_up);   results of a completed batch
#pragma once
#include <time.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bzcull.h"
#include "arena.h"

namespace bz {

inline uint64_t now_ns()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

enum ReqState : int { RQ_QUEUED, RQ_INFLIGHT, RQ_DONE };

template <class Dev>
struct SchedReq {
    uint64_t ticket = 0;
    bz_request r{};
    uint32_t n_objects = 0;
    uint64_t work_off = Arena::kFail, work_bytes = 0, out_off = Arena::kFail, out_bytes = 0;
    bz_result res{};
    int state = RQ_QUEUED;
    bool cancel = false;
    typename Dev::ReqData dev{};
};

template <class Dev>
class Scheduler {
public:
    using Req = SchedReq<Dev>;
    using ReqP = std::shared_ptr<Req>;
    static constexpr uint64_t kNone = Arena::kFail;

    struct Slot {
        std::vector<ReqP> batch;
        uint64_t batch_id = 0;
        bool busy = false;      // launching or in flight
        bool inflight = false;  // launched: the device owns the batch until poll reports it done
    };

    Scheduler(Dev& dev, const bz_server_cfg& cfg, uint64_t max_batch_objects, uint64_t align)
        : dev_(dev), cfg_(cfg), max_batch_objects_(max_batch_objects)
    {
        arena_.reset(cfg.hbm_budget_bytes, align);
        stats_.budget_bytes = arena_.size();
        slots_.resize(cfg.max_inflight_batches);
        counts_.resize(cfg.max_batch_requests);
    }

    ~Scheduler() { stop(); }

    void start()
    {
        running_ = true;
        launcher_ = std::thread([this] { launcher_loop(); });
        completer_ = std::thread([this] { completer_loop(); });
    }

    // Cancels what was never admitted and waits for the batches in flight.
    void stop()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (!running_ && !launcher_.joinable() && !completer_.joinable()) return;
            running_ = false;
            for (ReqP& q : queue_) {
                q->res.status = BZ_E_CANCELLED;
                q->state = RQ_DONE;
            }
            queue_.clear();
        }
        cv_launch_.notify_all();
        cv_done_.notify_all();
        if (launcher_.joinable()) launcher_.join();
        if (completer_.joinable()) completer_.join();  // drains the batches in flight first
    }

    bool running()
    {
        std::lock_guard<std::mutex> g(mu_);
        return running_;
    }

    // Queues a prepared request (sizes and device data filled in). arrival_ns: when it was due (0 = now).
    bz_status submit(const ReqP& q, uint64_t arrival_ns, uint64_t* ticket)
    {
        q->res.user_tag = q->r.user_tag;
        q->res.submit_ns = now_ns();
        q->res.arrival_ns = arrival_ns ? arrival_ns : q->res.submit_ns;
        std::lock_guard<std::mutex> g(mu_);
        if (!running_) return BZ_E_STATE;
        if (!arena_.fits_empty(q->work_bytes + q->out_bytes + arena_align())) {
            ++stats_.rejected;
            return BZ_E_BUDGET;
        }
        q->ticket = next_ticket_++;
        q->res.status = BZ_E_NOT_READY;
        reqs_[q->ticket] = q;
        queue_.push_back(q);
        ++stats_.submitted;
        stats_.queue_depth = (uint32_t)queue_.size();
        *ticket = q->ticket;
        cv_launch_.notify_one();
        return BZ_OK;
    }

    bz_status poll(uint64_t ticket, bz_result* out)
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) return BZ_E_NOT_FOUND;
        if (it->second->state != RQ_DONE) return BZ_E_NOT_READY;
        *out = it->second->res;
        return BZ_OK;
    }

    bz_status wait(uint64_t ticket, uint64_t timeout_ns, bz_result* out)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) return BZ_E_NOT_FOUND;
        ReqP q = it->second;
        const auto pred = [&] { return q->state == RQ_DONE; };
        if (timeout_ns == 0) cv_done_.wait(lk, pred);
        else if (!cv_done_.wait_for(lk, std::chrono::nanoseconds(timeout_ns), pred)) return BZ_E_TIMEOUT;
        *out = q->res;
        return BZ_OK;
    }

    // Removes a finished request from the table and hands it out (its output region still allocated).
    bz_status take(uint64_t ticket, ReqP& q)
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) return BZ_E_NOT_FOUND;
        if (it->second->state != RQ_DONE) return BZ_E_NOT_READY;
        q = it->second;
        reqs_.erase(it);
        return BZ_OK;
    }

    // Returns a taken request's regions to the budget.
    void release(Req& q)
    {
        std::lock_guard<std::mutex> g(mu_);
        release_locked(q);
        stats_.budget_in_use = arena_.used();
        cv_launch_.notify_all();
    }

    bz_status cancel(uint64_t ticket)
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) return BZ_E_NOT_FOUND;
        Req& q = *it->second;
        if (q.state == RQ_DONE) return BZ_E_STATE;
        q.cancel = true;
        if (q.state == RQ_QUEUED) {
            auto qi = std::find(queue_.begin(), queue_.end(), it->second);
            if (qi != queue_.end()) queue_.erase(qi);
            q.res.status = BZ_E_CANCELLED;
            q.res.complete_ns = now_ns();
            q.state = RQ_DONE;
            ++stats_.cancelled;
            stats_.queue_depth = (uint32_t)queue_.size();
            cv_done_.notify_all();
        }
        return BZ_OK;
    }

    bz_server_stats stats()
    {
        std::lock_guard<std::mutex> g(mu_);
        bz_server_stats s = stats_;
        s.budget_in_use = arena_.used();
        s.budget_peak = arena_.peak();
        return s;
    }

    const bz_server_cfg& cfg() const { return cfg_; }

private:
    uint64_t arena_align() const { return 256; }

    // Caller holds mu_.
    void release_locked(Req& q)
    {
        if (q.work_off != kNone) {
            arena_.release(q.work_off, q.work_bytes);
            q.work_off = kNone;
        }
        if (q.out_off != kNone) {
            arena_.release(q.out_off, q.out_bytes);
            q.out_off = kNone;
        }
        ++epoch_;
    }

    // Takes queued requests in arrival order while the budget admits them. Caller holds mu_.
    bool form_batch(std::vector<ReqP>& out)
    {
        const uint64_t t = now_ns();
        uint64_t batch_objects = 0;
        while (!queue_.empty() && out.size() < cfg_.max_batch_requests) {
            ReqP q = queue_.front();
            if (q->cancel || (q->r.deadline_ns && t > q->r.deadline_ns)) {
                queue_.pop_front();
                q->res.status = q->cancel ? BZ_E_CANCELLED : BZ_E_TIMEOUT;
                q->res.complete_ns = t;
                q->state = RQ_DONE;
                if (q->cancel) ++stats_.cancelled;
                else ++stats_.rejected;
                cv_done_.notify_all();
                continue;
            }
            if (!out.empty() && batch_objects + q->n_objects > max_batch_objects_) break;
            const uint64_t w = arena_.alloc(q->work_bytes);
            if (w == kNone) break;
            const uint64_t o = arena_.alloc(q->out_bytes);
            if (o == kNone) {
                arena_.release(w, q->work_bytes);
                break;
            }
            queue_.pop_front();
            q->work_off = w;
            q->out_off = o;
            q->res.admit_ns = t;
            q->state = RQ_INFLIGHT;
            batch_objects += q->n_objects;
            out.push_back(q);
        }
        stats_.budget_in_use = arena_.used();
        stats_.budget_peak = arena_.peak();
        stats_.queue_depth = (uint32_t)queue_.size();
        return !out.empty();
    }

    void launcher_loop()
    {
        dev_.thread_init();
        std::unique_lock<std::mutex> lk(mu_);
        uint64_t seen_epoch = ~0ull;
        while (running_) {
            uint32_t free_slot = ~0u;
            for (uint32_t k = 0; k < (uint32_t)slots_.size(); ++k)
                if (!slots_[k].busy) {
                    free_slot = k;
                    break;
                }
            std::vector<ReqP> batch;
            if (free_slot == ~0u || queue_.empty()) {
                cv_launch_.wait(lk);
                continue;
            }
            if (!form_batch(batch)) {
                // the head of the queue does not fit the budget now: wait until memory is released
                seen_epoch = epoch_;
                cv_launch_.wait(lk, [&] { return !running_ || epoch_ != seen_epoch || queue_.empty(); });
                continue;
            }
            Slot& s = slots_[free_slot];
            s.busy = true;
            s.batch = batch;
            s.batch_id = next_batch_++;
            ++stats_.batches;
            stats_.inflight_batches = 0;
            for (Slot& x : slots_) stats_.inflight_batches += x.busy ? 1u : 0u;
            lk.unlock();
            const bz_status st = dev_.launch(free_slot, s.batch);
            lk.lock();
            if (st) {
                finish(free_slot, st);  // complete the batch's requests with the error
            } else {
                s.inflight = true;
                cv_done_.notify_all();  // the completer polls the new batch
            }
        }
    }

    // Completes a slot's batch. Caller holds mu_.
    void finish(uint32_t slot, bz_status st)
    {
        Slot& s = slots_[slot];
        const uint64_t t = now_ns();
        const uint32_t B = (uint32_t)s.batch.size();
        bool gave_up = false;
        if (!st) dev_.collect(slot, B, counts_.data(), &gave_up);
        for (uint32_t k = 0; k < B; ++k) {
            Req& q = *s.batch[k];
            q.res.status = st ? st : (gave_up ? BZ_E_INTERNAL : BZ_OK);
            q.res.count = st ? 0u : counts_[k];
            q.res.complete_ns = t;
            q.res.batch_id = s.batch_id;
            q.res.batch_size = B;
            q.state = RQ_DONE;
            if (q.work_off != kNone) {
                arena_.release(q.work_off, q.work_bytes);
                q.work_off = kNone;
            }
            if (q.cancel || st || (cfg_.flags & BZ_SERVER_RELEASE_ON_COMPLETE)) {  // no fetch follows
                if (q.out_off != kNone) {
                    arena_.release(q.out_off, q.out_bytes);
                    q.out_off = kNone;
                }
                if (q.cancel) {
                    q.res.status = BZ_E_CANCELLED;
                    ++stats_.cancelled;
                }
            }
            if (!q.cancel) ++stats_.completed;
        }
        ++epoch_;
        s.batch.clear();
        s.busy = false;
        s.inflight = false;
        stats_.budget_in_use = arena_.used();
        stats_.inflight_batches = 0;
        for (Slot& x : slots_) stats_.inflight_batches += x.busy ? 1u : 0u;
        cv_done_.notify_all();
        cv_launch_.notify_all();
    }

    void completer_loop()
    {
        dev_.thread_init();
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            bool any = false;
            for (uint32_t k = 0; k < (uint32_t)slots_.size(); ++k) {
                if (!slots_[k].inflight) continue;
                any = true;
                lk.unlock();
                bool done = false;
                const bz_status st = dev_.poll(k, &done);
                lk.lock();
                if (st) finish(k, st);
                else if (done) finish(k, BZ_OK);
            }
            if (!any) {
                // a slot may be busy without being in flight yet: the launcher is launching its batch and marks it in
                // flight next (then notifies). Leaving before that would strand the batch's requests at stop.
                bool launching = false;
                for (const Slot& x : slots_) launching = launching || x.busy;
                if (!running_ && !launching) break;
                cv_done_.wait(lk);
                continue;
            }
            // batches in flight: poll again at once (a sleep would add the timer's granularity, tens of us, to every
            // request's latency)
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }

    Dev& dev_;
    bz_server_cfg cfg_;
    uint64_t max_batch_objects_;
    Arena arena_;
    std::mutex mu_;                      // queue, requests, slots' flags, arena, stats
    std::condition_variable cv_launch_;  // launcher: work or budget or a free slot
    std::condition_variable cv_done_;    // waiters and the completer
    std::deque<ReqP> queue_;
    std::unordered_map<uint64_t, ReqP> reqs_;
    std::vector<Slot> slots_;
    std::vector<uint32_t> counts_;       // collect() target (completer thread, under mu_)
    bool running_ = false;
    uint64_t next_ticket_ = 1, next_batch_ = 1, epoch_ = 0;
    bz_server_stats stats_{};
    std::thread launcher_, completer_;
};

}  // namespace bz
