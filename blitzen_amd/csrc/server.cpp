// server.cpp — the serving loop: cull requests over resident scenes with HBM-budget admission and continuous
// batching (a request scheduler the reference does not have; its frame loop culls one view per frame,
// vulkanDraw.cpp:1015-1033).
//
// Threads:
//   clients   bz_server_submit (any thread; a mutex-protected queue) and poll / wait / fetch / release / cancel.
//   launcher  takes queued requests in arrival order while the budget admits them (one arena region for the
//             request's depth buffer, pyramid and pyramid words, one for its commands), forms a batch of up to
//             max_batch_requests and launches it on a free stream slot: one batched depth raster, one batched
//             pyramid build, one batched cull (decide + emit) and one copy of the batch's draw counts to pinned
//             host memory, then an event. Up to max_inflight_batches batches run at once, so requests arriving
//             while others are on the GPU are admitted as soon as a slot and memory are free.
//   completer polls the in-flight batches' events, stamps completion times, frees the depth/pyramid regions and
//             wakes waiters. Command regions stay allocated until the client fetches or releases them.
// Each request's result is independent of what it was batched with: the batched kernels give every job the output
// of its own single-view cull.
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>

#include "arena.h"
#include "bz_ctx.h"

namespace bz {

namespace {

uint64_t now_ns()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

constexpr uint64_t kAlign = 256;
constexpr uint64_t kNone = Arena::kFail;
// Objects per batch: each slot's compaction scratch is allocated once for this many (plus one tile per request),
// so no allocation (and no device-wide synchronisation) happens while serving. A larger request runs alone.
constexpr uint64_t kBatchObjects = 64ull << 20;

enum ReqState : int { RQ_QUEUED, RQ_INFLIGHT, RQ_DONE };

struct Req {
    uint64_t ticket = 0;
    bz_request r{};
    SceneDev scene{};
    uint32_t n_objects = 0;
    PyrPlan plan{};
    uint64_t depth_bytes = 0, texel_bytes = 0, tail_bytes = 0;
    uint64_t work_off = kNone, work_bytes = 0, out_off = kNone, out_bytes = 0;
    bz_result res{};
    int state = RQ_QUEUED;
    bool cancel = false;
};
using ReqP = std::shared_ptr<Req>;

}  // namespace

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    CompactScratch cs{};
    CullJob* jobs_dev = nullptr;
    CullJob* jobs_host = nullptr;
    uint32_t jobs_cap = 0;
    PyrBuildJob* pjobs_dev = nullptr;
    PyrBuildJob* pjobs_host = nullptr;
    uint32_t pjobs_cap = 0;
    SynthJob* sjobs_dev = nullptr;
    SynthJob* sjobs_host = nullptr;
    uint32_t sjobs_cap = 0;
    uint32_t* counts_dev = nullptr;   // [max_batch] draw counts, then 4 error-flag words (err)
    uint32_t* counts_host = nullptr;  // pinned copy of the same
    uint32_t* err = nullptr;
    std::vector<ReqP> batch;
    uint64_t batch_id = 0;
    bool busy = false;                // launching or in flight
    bool inflight = false;            // launched: its event belongs to the current batch
};

struct Server {
    Ctx* ctx = nullptr;
    bz_server_cfg cfg{};
    char* base = nullptr;
    Arena arena;
    std::mutex mu;                      // queue, requests, slots' busy flags, arena, stats
    std::condition_variable cv_launch;  // launcher: work or budget or a free slot
    std::condition_variable cv_done;    // waiters and the completer
    std::deque<ReqP> queue;
    std::unordered_map<uint64_t, ReqP> reqs;
    std::vector<Slot> slots;
    bool running = false;
    uint64_t next_ticket = 1, next_batch = 1, epoch = 0;
    bz_server_stats stats{};
    hipStream_t fetch_st = nullptr;
    std::thread launcher, completer;

    ~Server();
    void launcher_loop();
    void completer_loop();
    bool form_batch(std::vector<ReqP>& out);
    bz_status launch(Slot& s, std::vector<ReqP>& batch);
    void finish(Slot& s, bz_status st);
    void release_locked(Req& q);
};

Ctx::Ctx() = default;

Ctx::~Ctx()
{
    if (server) {
        {
            std::lock_guard<std::mutex> g(server->mu);
            server->running = false;
        }
        server->cv_launch.notify_all();
        server->cv_done.notify_all();
        if (server->launcher.joinable()) server->launcher.join();
        if (server->completer.joinable()) server->completer.join();
        server.reset();
    }
}

Server::~Server()
{
    for (Slot& s : slots) {
        if (s.st) (void)hipStreamSynchronize(s.st);
        void* d[] = { s.cs.counts, s.cs.staging, s.jobs_dev, s.pjobs_dev, s.sjobs_dev, s.counts_dev };
        for (void* p : d)
            if (p) (void)hipFree(p);
        void* h[] = { s.jobs_host, s.pjobs_host, s.sjobs_host, s.counts_host };
        for (void* p : h)
            if (p) (void)hipHostFree(p);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.st) (void)hipStreamDestroy(s.st);
    }
    if (fetch_st) (void)hipStreamDestroy(fetch_st);
    if (base) (void)hipFree(base);
}

// Caller holds mu.
void Server::release_locked(Req& q)
{
    if (q.work_off != kNone) {
        arena.release(q.work_off, q.work_bytes);
        q.work_off = kNone;
    }
    if (q.out_off != kNone) {
        arena.release(q.out_off, q.out_bytes);
        q.out_off = kNone;
    }
    ++epoch;
}

// Takes queued requests in arrival order while the budget admits them. Caller holds mu.
bool Server::form_batch(std::vector<ReqP>& out)
{
    const uint64_t t = now_ns();
    uint64_t batch_objects = 0;
    while (!queue.empty() && out.size() < cfg.max_batch_requests) {
        ReqP q = queue.front();
        if (q->cancel || (q->r.deadline_ns && t > q->r.deadline_ns)) {
            queue.pop_front();
            q->res.status = q->cancel ? BZ_E_CANCELLED : BZ_E_TIMEOUT;
            q->res.complete_ns = t;
            q->state = RQ_DONE;
            if (q->cancel) ++stats.cancelled;
            else ++stats.rejected;
            cv_done.notify_all();
            continue;
        }
        if (!out.empty() && batch_objects + q->n_objects > kBatchObjects) break;
        const uint64_t w = arena.alloc(q->work_bytes);
        if (w == kNone) break;
        const uint64_t o = arena.alloc(q->out_bytes);
        if (o == kNone) {
            arena.release(w, q->work_bytes);
            break;
        }
        queue.pop_front();
        q->work_off = w;
        q->out_off = o;
        q->res.admit_ns = t;
        q->state = RQ_INFLIGHT;
        batch_objects += q->n_objects;
        out.push_back(q);
    }
    stats.budget_in_use = arena.used();
    stats.budget_peak = arena.peak();
    stats.queue_depth = (uint32_t)queue.size();
    return !out.empty();
}

bz_status Server::launch(Slot& s, std::vector<ReqP>& batch)
{
    const uint32_t B = (uint32_t)batch.size();
    const bz_tuning t = ctx->tune[1];
    const uint32_t tile = t.block * t.items;
    std::vector<SynthJob> sj(B);
    std::vector<uint32_t> sstart(B);
    std::vector<PyrBuildJob> pj;
    std::vector<uint32_t> pstart;
    std::vector<CullJob> cj;
    std::vector<uint32_t> cstart;
    uint32_t swg = 0, pwg = 0, tiles = 0;
    for (uint32_t k = 0; k < B; ++k) {
        Req& q = *batch[k];
        char* w = base + q.work_off;
        float* depth = reinterpret_cast<float*>(w);
        float* texels = reinterpret_cast<float*>(w + q.depth_bytes);
        auto* tail = reinterpret_cast<unsigned long long*>(w + q.depth_bytes + q.texel_bytes);
        const bool hiz = q.r.pass == BZ_PASS_TEMPORAL;
        SynthJob& S = sj[k];
        std::memset(&S, 0, sizeof S);
        S.seed = q.r.depth_seed;
        S.depth = depth;
        S.tail_zero = hiz ? tail : nullptr;
        S.n_tail = hiz ? q.plan.tiles_x * q.plan.tiles_y : 0u;
        S.w = q.r.depth_w;
        S.h = q.r.depth_h;
        S.occluders = q.r.occluders;
        S.tiles_x = (q.r.depth_w + kSynthTile - 1) / kSynthTile;
        S.p11 = q.r.view.proj5;
        S.znear = q.r.view.zNear;
        S.wg_start = swg;
        sstart[k] = swg;
        swg += S.tiles_x * ((q.r.depth_h + kSynthTile - 1) / kSynthTile);
        if (hiz) {
            PyrBuildJob P = q.plan.build_job(depth, q.r.depth_w, q.r.depth_h, texels, tail);
            P.wg_start = pwg;
            pstart.push_back(pwg);
            pwg += P.tiles_x * P.tiles_y;
            pj.push_back(P);
        }
        if (q.n_objects == 0) continue;  // the count stays 0 (cleared below)
        CullJob J;
        std::memset(&J, 0, sizeof J);
        bz_CameraViewData v = q.r.view;
        v.pyramidWidth = (float)q.plan.base_w;
        v.pyramidHeight = (float)q.plan.base_h;
        J.scene = q.scene;
        J.view = make_view_consts(v);
        if (hiz) J.pyr = q.plan.desc(texels);
        J.hiz = hiz ? 1u : 0u;
        J.vis_in = nullptr;  // PLAIN and TEMPORAL read and write no visibility
        J.vis_out = nullptr;
        J.cmds = reinterpret_cast<uint32_t*>(base + q.out_off);
        J.count = s.counts_dev + k;
        J.begin = 0;
        J.n = q.n_objects;
        J.capacity = q.n_objects;
        J.pass = q.r.pass;
        J.n_tiles = tiles_for(q.n_objects, t.block, t.items);
        J.tile_start = tiles;
        cstart.push_back(tiles);
        tiles += J.n_tiles;
        cj.push_back(J);
    }
    hipError_t e = hipSuccess;
    // compaction scratch of this slot: allocated at start for kBatchObjects; only a single request larger than
    // that (or a larger tile after a tuning change) grows it
    if (s.cs.capacity_tiles < tiles || s.cs.capacity_entries < (uint64_t)tiles * tile) {
        if (s.cs.counts) (void)hipFree(s.cs.counts);
        if (s.cs.staging) (void)hipFree(s.cs.staging);
        s.cs = CompactScratch{};
        const uint32_t ct = std::max<uint32_t>(tiles, 1u << 12);
        const uint64_t ce = std::max<uint64_t>((uint64_t)tiles * tile, 1ull << 20);
        e = hipMalloc((void**)&s.cs.counts, (size_t)ct * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.cs.staging, (size_t)ce * sizeof(uint2));
        if (e != hipSuccess) return ctx->hip_fail(e, "server compaction scratch");
        s.cs.capacity_tiles = ct;
        s.cs.capacity_entries = ce;
    }
    bz_status st = upload_table(ctx, s.sjobs_dev, s.sjobs_host, s.sjobs_cap, s.done, sj.data(), sstart.data(), B,
                                s.st, "server raster jobs");
    if (!st && !pj.empty())
        st = upload_table(ctx, s.pjobs_dev, s.pjobs_host, s.pjobs_cap, s.done, pj.data(), pstart.data(),
                          (uint32_t)pj.size(), s.st, "server pyramid jobs");
    if (!st && !cj.empty())
        st = upload_table(ctx, s.jobs_dev, s.jobs_host, s.jobs_cap, s.done, cj.data(), cstart.data(),
                          (uint32_t)cj.size(), s.st, "server cull jobs");
    if (st) return st;
    e = launch_fill_u32(s.counts_dev, 0u, B, s.st);
    if (e == hipSuccess)
        e = launch_synth_depth_batch(s.sjobs_dev, table_starts(s.sjobs_dev, s.sjobs_cap), B, swg, s.st);
    if (e == hipSuccess && !pj.empty())
        e = launch_pyramid_batch(s.pjobs_dev, table_starts(s.pjobs_dev, s.pjobs_cap), (uint32_t)pj.size(), pwg,
                                 s.err + 1, s.st);
    if (e == hipSuccess && !cj.empty())
        e = launch_cull_batch(s.jobs_dev, table_starts(s.jobs_dev, s.jobs_cap), (uint32_t)cj.size(), tiles, s.cs,
                              t.block, t.items, s.st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.counts_host, s.counts_dev, (size_t)B * 4, hipMemcpyDeviceToHost, s.st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.counts_host + cfg.max_batch_requests, s.err, 16, hipMemcpyDeviceToHost, s.st);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.st);
    if (e != hipSuccess) return ctx->hip_fail(e, "server batch launch");
    return BZ_OK;
}

void Server::launcher_loop()
{
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    uint64_t seen_epoch = ~0ull;
    while (running) {
        Slot* free_slot = nullptr;
        for (Slot& s : slots)
            if (!s.busy) {
                free_slot = &s;
                break;
            }
        std::vector<ReqP> batch;
        if (!free_slot || queue.empty()) {
            cv_launch.wait(lk);
            continue;
        }
        if (!form_batch(batch)) {
            // the head of the queue does not fit the budget now: wait until memory is released
            seen_epoch = epoch;
            cv_launch.wait(lk, [&] { return !running || epoch != seen_epoch || queue.empty(); });
            continue;
        }
        Slot& s = *free_slot;
        s.busy = true;
        s.batch = batch;
        s.batch_id = next_batch++;
        ++stats.batches;
        stats.inflight_batches = 0;
        for (Slot& x : slots) stats.inflight_batches += x.busy ? 1u : 0u;
        lk.unlock();
        const bz_status st = launch(s, s.batch);
        lk.lock();
        if (st) {
            finish(s, st);  // complete the batch's requests with the error
        } else {
            s.inflight = true;
            cv_done.notify_all();  // the completer polls the new batch
        }
    }
}

// Completes a slot's batch. Caller holds mu.
void Server::finish(Slot& s, bz_status st)
{
    const uint64_t t = now_ns();
    const uint32_t B = (uint32_t)s.batch.size();
    bool waited_out = false;
    if (!st) {
        const uint32_t* flags = s.counts_host + cfg.max_batch_requests;
        if (flags[0] || flags[1]) {
            waited_out = true;
            (void)launch_fill_u32(s.err, 0u, 4, s.st);  // ordered before the slot's next batch
        }
    }
    for (uint32_t k = 0; k < B; ++k) {
        Req& q = *s.batch[k];
        q.res.status = st ? st : (waited_out ? BZ_E_INTERNAL : BZ_OK);
        q.res.count = st ? 0u : s.counts_host[k];
        q.res.complete_ns = t;
        q.res.batch_id = s.batch_id;
        q.res.batch_size = B;
        q.state = RQ_DONE;
        if (q.work_off != kNone) {
            arena.release(q.work_off, q.work_bytes);
            q.work_off = kNone;
        }
        if (q.cancel || st || (cfg.flags & BZ_SERVER_RELEASE_ON_COMPLETE)) {  // no fetch follows
            if (q.out_off != kNone) {
                arena.release(q.out_off, q.out_bytes);
                q.out_off = kNone;
            }
            if (q.cancel) {
                q.res.status = BZ_E_CANCELLED;
                ++stats.cancelled;
            }
        }
        if (!q.cancel) ++stats.completed;
    }
    ++epoch;
    s.batch.clear();
    s.busy = false;
    s.inflight = false;
    stats.budget_in_use = arena.used();
    stats.inflight_batches = 0;
    for (Slot& x : slots) stats.inflight_batches += x.busy ? 1u : 0u;
    cv_done.notify_all();
    cv_launch.notify_all();
}

void Server::completer_loop()
{
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        bool any = false;
        for (Slot& s : slots) {
            if (!s.inflight) continue;
            any = true;
            hipEvent_t ev = s.done;
            lk.unlock();
            const hipError_t q = hipEventQuery(ev);
            lk.lock();
            if (q == hipSuccess) finish(s, BZ_OK);
            else if (q != hipErrorNotReady) finish(s, ctx->hip_fail(q, "server batch"));
        }
        if (!any) {
            if (!running) break;
            cv_done.wait(lk);
            continue;
        }
        // batches in flight: poll again at once (a sleep would add the timer's granularity, tens of us, to every
        // request's latency)
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
    }
}

}  // namespace bz

using namespace bz;

namespace {

Server* running_server(bz_ctx* ctx)
{
    return (ctx && ctx->server && ctx->server->running) ? ctx->server.get() : nullptr;
}

bz_status submit_one(bz_ctx* ctx, Server* sv, const bz_request* r, uint64_t* ticket)
{
    if (r->pass != BZ_PASS_TEMPORAL && r->pass != BZ_PASS_PLAIN)
        return ctx->fail(BZ_E_INVALID, "server requests use BZ_PASS_TEMPORAL or BZ_PASS_PLAIN");
    if (r->depth_w < 2 || r->depth_h < 2 || r->depth_w > 16384 || r->depth_h > 16384)
        return ctx->fail(BZ_E_INVALID, "depth extent must be 2..16384 on each axis");
    if (r->occluders > kSynthMaxRects) return ctx->fail(BZ_E_INVALID, "at most 4096 occluders");
    auto q = std::make_shared<Req>();
    q->r = *r;
    {
        Scene* s = ctx->scene(r->scene);
        if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
        q->scene = s->dev();
        q->n_objects = s->n_objects;
    }
    uint32_t bw = 0, bh = 0, mips = 0;
    bz_pyramid_reference_extent(r->depth_w, r->depth_h, &bw, &bh, &mips);
    if (mips == 0 || !plan_pyramid(bw, bh, std::min<uint32_t>(mips, kMaxMips), &q->plan) || !q->plan.single_pass)
        return ctx->fail(BZ_E_INVALID, "depth extent gives an unsupported pyramid");
    q->depth_bytes = Arena::round_up((uint64_t)r->depth_w * r->depth_h * 4, kAlign);
    q->texel_bytes = Arena::round_up(q->plan.n_texels * 4, kAlign);
    q->tail_bytes = Arena::round_up((uint64_t)q->plan.tiles_x * q->plan.tiles_y * 8, kAlign);
    q->work_bytes = q->depth_bytes + q->texel_bytes + q->tail_bytes;
    q->out_bytes = Arena::round_up((uint64_t)std::max(q->n_objects, 1u) * 24, kAlign);
    q->res.user_tag = r->user_tag;
    q->res.submit_ns = now_ns();
    std::lock_guard<std::mutex> g(sv->mu);
    if (!sv->arena.fits_empty(q->work_bytes + q->out_bytes + kAlign)) {
        ++sv->stats.rejected;
        return ctx->fail(BZ_E_BUDGET, "the request does not fit the server's HBM budget");
    }
    q->ticket = sv->next_ticket++;
    q->res.status = BZ_E_NOT_READY;
    sv->reqs[q->ticket] = q;
    sv->queue.push_back(q);
    ++sv->stats.submitted;
    sv->stats.queue_depth = (uint32_t)sv->queue.size();
    *ticket = q->ticket;
    sv->cv_launch.notify_one();
    return BZ_OK;
}

}  // namespace

extern "C" {

bz_status bz_server_start(bz_ctx* ctx, const bz_server_cfg* cfg)
{
    if (!ctx || !cfg) return BZ_E_INVALID;
    if (ctx->server) return ctx->fail(BZ_E_STATE, "the server is already running");
    if (cfg->hbm_budget_bytes < (1ull << 20)) return ctx->fail(BZ_E_INVALID, "budget below 1 MiB");
    if (cfg->flags & ~BZ_SERVER_RELEASE_ON_COMPLETE) return ctx->fail(BZ_E_INVALID, "unknown server flags");
    if (cfg->max_inflight_batches == 0 || cfg->max_inflight_batches > 16)
        return ctx->fail(BZ_E_INVALID, "max_inflight_batches must be 1..16");
    if (cfg->max_batch_requests == 0 || cfg->max_batch_requests > 1024)
        return ctx->fail(BZ_E_INVALID, "max_batch_requests must be 1..1024");
    std::unique_ptr<Server> sv(new (std::nothrow) Server());
    if (!sv) return BZ_E_OOM;
    sv->ctx = ctx;
    sv->cfg = *cfg;
    hipError_t e = hipMalloc((void**)&sv->base, cfg->hbm_budget_bytes);
    if (e != hipSuccess) return ctx->hip_fail(e, "server arena");
    sv->arena.reset(cfg->hbm_budget_bytes, kAlign);
    sv->stats.budget_bytes = sv->arena.size();
    sv->slots.resize(cfg->max_inflight_batches);
    for (Slot& s : sv->slots) {
        e = hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        const size_t words = (size_t)cfg->max_batch_requests + 4;
        if (e == hipSuccess) e = hipMalloc((void**)&s.counts_dev, words * 4);
        if (e == hipSuccess) e = hipMemset(s.counts_dev, 0, words * 4);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.counts_host, words * 4, hipHostMallocDefault);
        if (e == hipSuccess) std::memset(s.counts_host, 0, words * 4);
        s.err = s.counts_dev + cfg->max_batch_requests;
        const uint64_t tiles = kBatchObjects / 256 + cfg->max_batch_requests;  // smallest tile: 256 objects
        const uint64_t entries = kBatchObjects + (uint64_t)cfg->max_batch_requests * 2048;
        if (e == hipSuccess) e = hipMalloc((void**)&s.cs.counts, (size_t)tiles * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.cs.staging, (size_t)entries * sizeof(uint2));
        s.cs.capacity_tiles = (uint32_t)tiles;
        s.cs.capacity_entries = entries;
        if (e != hipSuccess) return ctx->hip_fail(e, "server slot");
    }
    e = hipStreamCreateWithFlags(&sv->fetch_st, hipStreamNonBlocking);
    if (e != hipSuccess) return ctx->hip_fail(e, "server fetch stream");
    sv->running = true;
    ctx->server = std::move(sv);
    Server* s = ctx->server.get();
    s->launcher = std::thread([s] { s->launcher_loop(); });
    s->completer = std::thread([s] { s->completer_loop(); });
    return BZ_OK;
}

bz_status bz_server_stop(bz_ctx* ctx)
{
    if (!ctx) return BZ_E_INVALID;
    if (!ctx->server) return BZ_OK;
    Server* s = ctx->server.get();
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->running = false;
        for (ReqP& q : s->queue) {  // never admitted
            q->res.status = BZ_E_CANCELLED;
            q->state = RQ_DONE;
        }
        s->queue.clear();
    }
    s->cv_launch.notify_all();
    s->cv_done.notify_all();
    if (s->launcher.joinable()) s->launcher.join();
    if (s->completer.joinable()) s->completer.join();  // drains the batches in flight first
    ctx->server.reset();
    return BZ_OK;
}

bz_status bz_server_submit(bz_ctx* ctx, const bz_request* req, uint64_t* ticket)
{
    if (!ctx || !req || !ticket) return BZ_E_INVALID;
    *ticket = 0;
    Server* sv = running_server(ctx);
    if (!sv) return ctx->fail(BZ_E_STATE, "the server is not running");
    return submit_one(ctx, sv, req, ticket);
}

bz_status bz_server_submit_schedule(bz_ctx* ctx, const bz_request* reqs, const uint64_t* offsets_ns, uint32_t n,
                                    uint64_t* tickets, bz_status* statuses)
{
    if (!ctx || (n && (!reqs || !offsets_ns || !tickets))) return BZ_E_INVALID;
    Server* sv = running_server(ctx);
    if (!sv) return ctx->fail(BZ_E_STATE, "the server is not running");
    const uint64_t t0 = now_ns();
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t due = t0 + offsets_ns[i];
        uint64_t t = now_ns();
        if (due > t + 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(due - t - 100000));
        while ((t = now_ns()) < due) {
        }
        tickets[i] = 0;
        const bz_status s = submit_one(ctx, sv, &reqs[i], &tickets[i]);
        if (statuses) statuses[i] = s;
    }
    return BZ_OK;
}

bz_status bz_server_poll(bz_ctx* ctx, uint64_t ticket, bz_result* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    Server* s = ctx->server.get();
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BZ_E_NOT_FOUND;
    if (it->second->state != RQ_DONE) return BZ_E_NOT_READY;
    *out = it->second->res;
    return BZ_OK;
}

bz_status bz_server_wait(bz_ctx* ctx, uint64_t ticket, uint64_t timeout_ns, bz_result* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    Server* s = ctx->server.get();
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BZ_E_NOT_FOUND;
    ReqP q = it->second;
    const auto pred = [&] { return q->state == RQ_DONE; };
    if (timeout_ns == 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::nanoseconds(timeout_ns), pred)) return BZ_E_TIMEOUT;
    *out = q->res;
    return BZ_OK;
}

static bz_status take_done(Server* s, uint64_t ticket, std::shared_ptr<Req>& q)
{
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BZ_E_NOT_FOUND;
    if (it->second->state != RQ_DONE) return BZ_E_NOT_READY;
    q = it->second;
    s->reqs.erase(it);
    return BZ_OK;
}

bz_status bz_server_fetch(bz_ctx* ctx, uint64_t ticket, bz_IndirectDrawData* cmds, uint32_t capacity, uint32_t* count)
{
    if (!ctx || (capacity && !cmds)) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    Server* s = ctx->server.get();
    std::shared_ptr<Req> q;
    uint64_t off = kNone;
    {
        std::lock_guard<std::mutex> g(s->mu);
        bz_status st = take_done(s, ticket, q);
        if (st) return st;
        off = q->out_off;
    }
    bz_status st = q->res.status;
    const uint32_t n = std::min(q->res.count, capacity);
    if (!st && n && off == kNone) st = ctx->fail(BZ_E_STATE, "the commands were released at completion");
    if (!st && n && off != kNone) {
        hipError_t e = hipMemcpyAsync(cmds, s->base + off, (size_t)n * sizeof(bz_IndirectDrawData),
                                      hipMemcpyDeviceToHost, s->fetch_st);
        if (e == hipSuccess) e = hipStreamSynchronize(s->fetch_st);
        if (e != hipSuccess) st = ctx->hip_fail(e, "server fetch");
    }
    if (count) *count = q->res.count;
    std::lock_guard<std::mutex> g(s->mu);
    s->release_locked(*q);
    s->stats.budget_in_use = s->arena.used();
    s->cv_launch.notify_all();
    return st;
}

bz_status bz_server_release(bz_ctx* ctx, uint64_t ticket)
{
    if (!ctx) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    Server* s = ctx->server.get();
    std::lock_guard<std::mutex> g(s->mu);
    std::shared_ptr<Req> q;
    bz_status st = take_done(s, ticket, q);
    if (st) return st;
    s->release_locked(*q);
    s->stats.budget_in_use = s->arena.used();
    s->cv_launch.notify_all();
    return BZ_OK;
}

bz_status bz_server_cancel(bz_ctx* ctx, uint64_t ticket)
{
    if (!ctx) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    Server* s = ctx->server.get();
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BZ_E_NOT_FOUND;
    Req& q = *it->second;
    if (q.state == RQ_DONE) return ctx->fail(BZ_E_STATE, "the request has completed");
    q.cancel = true;
    if (q.state == RQ_QUEUED) {
        auto qi = std::find(s->queue.begin(), s->queue.end(), it->second);
        if (qi != s->queue.end()) s->queue.erase(qi);
        q.res.status = BZ_E_CANCELLED;
        q.state = RQ_DONE;
        ++s->stats.cancelled;
        s->stats.queue_depth = (uint32_t)s->queue.size();
        s->cv_done.notify_all();
    }
    return BZ_OK;
}

bz_status bz_server_get_stats(bz_ctx* ctx, bz_server_stats* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    Server* s = ctx->server.get();
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    out->budget_in_use = s->arena.used();
    out->budget_peak = s->arena.peak();
    return BZ_OK;
}

}  // extern "C"
