// server.cpp — the serving loop over resident scenes: the HIP device of the scheduler in server_core.h (a request
// scheduler the reference does not have; its frame loop culls one view per frame, vulkanDraw.cpp:1015-1033).
//
// A batch on a slot is one batched depth raster, one batched pyramid build, one batched cull (decide + emit) and one
// copy of the batch's draw counts to pinned host memory, then an event the completer polls. Each request's result
// is independent of what it was batched with: the batched kernels give every job the output of its own single-view
// cull.
#include <algorithm>
#include <cstring>

#include "bz_ctx.h"
#include "server_core.h"

namespace bz {

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint64_t kNone = Arena::kFail;
// Objects per batch: each slot's compaction scratch is allocated once for this many (plus one tile per request),
// so no allocation (and no device-wide synchronisation) happens while serving. A larger request runs alone.
constexpr uint64_t kBatchObjects = 64ull << 20;

}  // namespace

struct HipDevice;
using Req = SchedReq<HipDevice>;
using ReqP = std::shared_ptr<Req>;

struct HipSlot {
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
};

struct HipDevice {
    struct ReqData {
        SceneDev scene{};
        PyrPlan plan{};
        uint64_t depth_bytes = 0, texel_bytes = 0, tail_bytes = 0;
    };
    Ctx* ctx = nullptr;
    char* base = nullptr;             // the arena: one allocation of the HBM budget
    uint32_t max_batch = 0;
    std::vector<HipSlot> slots;
    hipStream_t fetch_st = nullptr;

    bz_status init(const bz_server_cfg& cfg);
    ~HipDevice();
    void thread_init() { (void)hipSetDevice(ctx->device); }
    bz_status launch(uint32_t slot, const std::vector<ReqP>& batch);
    bz_status poll(uint32_t slot, bool* done)
    {
        const hipError_t q = hipEventQuery(slots[slot].done);
        *done = q == hipSuccess;
        return (q == hipSuccess || q == hipErrorNotReady) ? BZ_OK : ctx->hip_fail(q, "server batch");
    }
    void collect(uint32_t slot, uint32_t n, uint32_t* counts, bool* gave_up)
    {
        HipSlot& s = slots[slot];
        std::memcpy(counts, s.counts_host, (size_t)n * 4);
        const uint32_t* flags = s.counts_host + max_batch;
        *gave_up = flags[0] || flags[1];
        if (*gave_up) (void)launch_fill_u32(s.err, 0u, 4, s.st);  // ordered before the slot's next batch
    }
};

struct Server {
    HipDevice dev;
    std::unique_ptr<Scheduler<HipDevice>> sched;
};

bz_status HipDevice::init(const bz_server_cfg& cfg)
{
    max_batch = cfg.max_batch_requests;
    hipError_t e = hipMalloc((void**)&base, cfg.hbm_budget_bytes);
    if (e != hipSuccess) return ctx->hip_fail(e, "server arena");
    slots.resize(cfg.max_inflight_batches);
    for (HipSlot& s : slots) {
        e = hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        const size_t words = (size_t)cfg.max_batch_requests + 4;
        if (e == hipSuccess) e = hipMalloc((void**)&s.counts_dev, words * 4);
        if (e == hipSuccess) e = hipMemset(s.counts_dev, 0, words * 4);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.counts_host, words * 4, hipHostMallocDefault);
        if (e == hipSuccess) std::memset(s.counts_host, 0, words * 4);
        s.err = s.counts_dev + cfg.max_batch_requests;
        const uint64_t tiles = kBatchObjects / 256 + cfg.max_batch_requests;  // smallest tile: 256 objects
        const uint64_t entries = kBatchObjects + (uint64_t)cfg.max_batch_requests * 2048;
        if (e == hipSuccess) e = hipMalloc((void**)&s.cs.counts, (size_t)counts_words((uint32_t)tiles) * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.cs.staging, (size_t)entries * sizeof(uint2));
        s.cs.capacity_tiles = (uint32_t)tiles;
        s.cs.capacity_entries = entries;
        set_prefix(s.cs);
        if (e != hipSuccess) return ctx->hip_fail(e, "server slot");
    }
    e = hipStreamCreateWithFlags(&fetch_st, hipStreamNonBlocking);
    if (e != hipSuccess) return ctx->hip_fail(e, "server fetch stream");
    return BZ_OK;
}

HipDevice::~HipDevice()
{
    for (HipSlot& s : slots) {
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

Ctx::Ctx() = default;

Ctx::~Ctx()
{
    if (server) {
        server->sched.reset();  // stops the threads and drains the batches in flight
        server.reset();
    }
}

bz_status HipDevice::launch(uint32_t slot, const std::vector<ReqP>& batch)
{
    HipSlot& s = slots[slot];
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
        const Req& q = *batch[k];
        char* w = base + q.work_off;
        float* depth = reinterpret_cast<float*>(w);
        float* texels = reinterpret_cast<float*>(w + q.dev.depth_bytes);
        auto* tail = reinterpret_cast<unsigned long long*>(w + q.dev.depth_bytes + q.dev.texel_bytes);
        const bool hiz = q.r.pass == BZ_PASS_TEMPORAL;
        SynthJob& S = sj[k];
        std::memset(&S, 0, sizeof S);
        S.seed = q.r.depth_seed;
        S.depth = depth;
        S.tail_zero = hiz ? tail : nullptr;
        S.n_tail = hiz ? q.dev.plan.tiles_x * q.dev.plan.tiles_y : 0u;
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
            PyrBuildJob P = q.dev.plan.build_job(depth, q.r.depth_w, q.r.depth_h, texels, tail);
            P.wg_start = pwg;
            pstart.push_back(pwg);
            pwg += P.tiles_x * P.tiles_y;
            pj.push_back(P);
        }
        if (q.n_objects == 0) continue;  // the count stays 0 (cleared below)
        CullJob J;
        std::memset(&J, 0, sizeof J);
        bz_CameraViewData v = q.r.view;
        v.pyramidWidth = (float)q.dev.plan.base_w;
        v.pyramidHeight = (float)q.dev.plan.base_h;
        J.scene = q.dev.scene;
        J.view = make_view_consts(v);
        if (hiz) J.pyr = q.dev.plan.desc(texels);
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
        e = hipMalloc((void**)&s.cs.counts, (size_t)counts_words(ct) * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.cs.staging, (size_t)ce * sizeof(uint2));
        if (e != hipSuccess) return ctx->hip_fail(e, "server compaction scratch");
        s.cs.capacity_tiles = ct;
        s.cs.capacity_entries = ce;
        set_prefix(s.cs);
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
    {
        uint32_t mt = 0;
        for (auto& q : pj) mt = std::max<uint32_t>(mt, q.tiles_x * q.tiles_y);
        e = launch_pyramid_batch(s.pjobs_dev, (uint32_t)pj.size(), mt, s.err + 1, s.st);
    }
    if (e == hipSuccess && !cj.empty())
        e = launch_cull_batch(cj.data(), s.jobs_dev, table_starts(s.jobs_dev, s.jobs_cap), (uint32_t)cj.size(),
                              tiles, max_job_tiles(cj.data(), cj.size()), s.cs, t.block, t.items, s.st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.counts_host, s.counts_dev, (size_t)B * 4, hipMemcpyDeviceToHost, s.st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.counts_host + max_batch, s.err, 16, hipMemcpyDeviceToHost, s.st);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.st);
    if (e != hipSuccess) return ctx->hip_fail(e, "server batch launch");
    return BZ_OK;
}

}  // namespace bz

using namespace bz;

namespace {

Scheduler<HipDevice>* running_sched(bz_ctx* ctx)
{
    return (ctx && ctx->server && ctx->server->sched && ctx->server->sched->running()) ? ctx->server->sched.get()
                                                                                     : nullptr;
}

// arrival_ns: when the request was due (0: now, at submission)
bz_status submit_one(bz_ctx* ctx, Scheduler<HipDevice>* sv, const bz_request* r, uint64_t* ticket,
                     uint64_t arrival_ns = 0)
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
        q->dev.scene = s->dev();
        q->n_objects = s->n_objects;
    }
    uint32_t bw = 0, bh = 0, mips = 0;
    bz_pyramid_reference_extent(r->depth_w, r->depth_h, &bw, &bh, &mips);
    if (mips == 0 || !plan_pyramid(bw, bh, std::min<uint32_t>(mips, kMaxMips), &q->dev.plan) ||
        !q->dev.plan.single_pass)
        return ctx->fail(BZ_E_INVALID, "depth extent gives an unsupported pyramid");
    q->dev.depth_bytes = Arena::round_up((uint64_t)r->depth_w * r->depth_h * 4, kAlign);
    q->dev.texel_bytes = Arena::round_up((q->dev.plan.n_texels + 1) * 4, kAlign);  // + the Hi-Z pairs' spare texel
    q->dev.tail_bytes = Arena::round_up((uint64_t)q->dev.plan.tiles_x * q->dev.plan.tiles_y * 8, kAlign);
    q->work_bytes = q->dev.depth_bytes + q->dev.texel_bytes + q->dev.tail_bytes;
    q->out_bytes = Arena::round_up((uint64_t)std::max(q->n_objects, 1u) * 24, kAlign);
    const bz_status st = sv->submit(q, arrival_ns, ticket);
    if (st == BZ_E_BUDGET) return ctx->fail(st, "the request does not fit the server's HBM budget");
    if (st) return ctx->fail(st, "the server is not running");
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
    sv->dev.ctx = ctx;
    const bz_status st = sv->dev.init(*cfg);
    if (st) return st;
    sv->sched.reset(new (std::nothrow) Scheduler<HipDevice>(sv->dev, *cfg, kBatchObjects, kAlign));
    if (!sv->sched) return BZ_E_OOM;
    sv->sched->start();
    ctx->server = std::move(sv);
    return BZ_OK;
}

bz_status bz_server_stop(bz_ctx* ctx)
{
    if (!ctx) return BZ_E_INVALID;
    if (!ctx->server) return BZ_OK;
    ctx->server->sched.reset();  // cancels what was never admitted, drains the batches in flight
    ctx->server.reset();
    return BZ_OK;
}

bz_status bz_server_submit(bz_ctx* ctx, const bz_request* req, uint64_t* ticket)
{
    if (!ctx || !req || !ticket) return BZ_E_INVALID;
    *ticket = 0;
    Scheduler<HipDevice>* sv = running_sched(ctx);
    if (!sv) return ctx->fail(BZ_E_STATE, "the server is not running");
    return submit_one(ctx, sv, req, ticket);
}

bz_status bz_server_submit_schedule(bz_ctx* ctx, const bz_request* reqs, const uint64_t* offsets_ns, uint32_t n,
                                    uint64_t* tickets, bz_status* statuses)
{
    if (!ctx || (n && (!reqs || !offsets_ns || !tickets))) return BZ_E_INVALID;
    Scheduler<HipDevice>* sv = running_sched(ctx);
    if (!sv) return ctx->fail(BZ_E_STATE, "the server is not running");
    const uint64_t t0 = now_ns();
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t due = t0 + offsets_ns[i];
        uint64_t t = now_ns();
        if (due > t + 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(due - t - 100000));
        while ((t = now_ns()) < due) {
        }
        tickets[i] = 0;
        const bz_status s = submit_one(ctx, sv, &reqs[i], &tickets[i], due);
        if (statuses) statuses[i] = s;
    }
    return BZ_OK;
}

bz_status bz_server_poll(bz_ctx* ctx, uint64_t ticket, bz_result* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    return ctx->server->sched->poll(ticket, out);
}

bz_status bz_server_wait(bz_ctx* ctx, uint64_t ticket, uint64_t timeout_ns, bz_result* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    return ctx->server->sched->wait(ticket, timeout_ns, out);
}

bz_status bz_server_fetch(bz_ctx* ctx, uint64_t ticket, bz_IndirectDrawData* cmds, uint32_t capacity, uint32_t* count)
{
    if (!ctx || (capacity && !cmds)) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    Server* s = ctx->server.get();
    ReqP q;
    bz_status st = s->sched->take(ticket, q);
    if (st) return st;
    st = q->res.status;
    const uint64_t off = q->out_off;
    const uint32_t n = std::min(q->res.count, capacity);
    if (!st && n && off == kNone) st = ctx->fail(BZ_E_STATE, "the commands were released at completion");
    if (!st && n && off != kNone) {
        hipError_t e = hipMemcpyAsync(cmds, s->dev.base + off, (size_t)n * sizeof(bz_IndirectDrawData),
                                      hipMemcpyDeviceToHost, s->dev.fetch_st);
        if (e == hipSuccess) e = hipStreamSynchronize(s->dev.fetch_st);
        if (e != hipSuccess) st = ctx->hip_fail(e, "server fetch");
    }
    if (count) *count = q->res.count;
    s->sched->release(*q);
    return st;
}

bz_status bz_server_release(bz_ctx* ctx, uint64_t ticket)
{
    if (!ctx) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    ReqP q;
    const bz_status st = ctx->server->sched->take(ticket, q);
    if (st) return st;
    ctx->server->sched->release(*q);
    return BZ_OK;
}

bz_status bz_server_cancel(bz_ctx* ctx, uint64_t ticket)
{
    if (!ctx) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    const bz_status st = ctx->server->sched->cancel(ticket);
    if (st == BZ_E_STATE) return ctx->fail(BZ_E_STATE, "the request has completed");
    return st;
}

bz_status bz_server_get_stats(bz_ctx* ctx, bz_server_stats* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    if (!ctx->server) return BZ_E_STATE;
    *out = ctx->server->sched->stats();
    return BZ_OK;
}

}  // extern "C"
