// autotune.cpp — picks the workgroup size, objects per thread and persistent grid of the cull kernels by timing
// every candidate on the caller's own jobs (every candidate gives identical results; only the time differs).
// Results can be cached in a small JSON file keyed by device name, kernel and object-count bucket, one entry per
// line:   "<device>|k<kernel>|b<floor(log2(objects))>": {"block": B, "items": I, "grid_per_cu": G, "us": T},
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "bz_ctx.h"

namespace bz {
namespace {

struct CacheEntry {
    bz_tuning t;
    double us;
};

std::map<std::string, CacheEntry> read_cache(const char* path)
{
    std::map<std::string, CacheEntry> m;
    FILE* f = std::fopen(path, "r");
    if (!f) return m;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char key[256];
        CacheEntry e{};
        if (std::sscanf(line, " \"%255[^\"]\": {\"block\": %u, \"items\": %u, \"grid_per_cu\": %u, \"us\": %lf}", key,
                        &e.t.block, &e.t.items, &e.t.grid_per_cu, &e.us) == 5)
            m[key] = e;
    }
    std::fclose(f);
    return m;
}

bool write_cache(const char* path, const std::map<std::string, CacheEntry>& m)
{
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "{\n");
    size_t k = 0;
    for (const auto& kv : m) {
        std::fprintf(f, "  \"%s\": {\"block\": %u, \"items\": %u, \"grid_per_cu\": %u, \"us\": %.3f}%s\n",
                     kv.first.c_str(), kv.second.t.block, kv.second.t.items, kv.second.t.grid_per_cu, kv.second.us,
                     ++k < m.size() ? "," : "");
    }
    std::fprintf(f, "}\n");
    const bool ok = std::fclose(f) == 0;
    return ok && std::rename(tmp.c_str(), path) == 0;
}

bool valid_tuning(const bz_tuning& t)
{
    return (t.block == 256 || t.block == 512) && (t.items == 2 || t.items == 4) && t.grid_per_cu <= 64;
}

}  // namespace
}  // namespace bz

using namespace bz;

extern "C" bz_status bz_autotune(bz_ctx* ctx, uint32_t kernel, const bz_view_job* jobs, uint32_t n_jobs,
                                 const char* cache_path, bz_tuning* best, void* stream)
{
    if (!ctx || !best) return BZ_E_INVALID;
    if (kernel > 1) return ctx->fail(BZ_E_INVALID, "kernel must be 0 (single view) or 1 (batched)");
    if (!jobs || n_jobs == 0) return ctx->fail(BZ_E_INVALID, "autotune needs at least one job");
    hipStream_t st = (hipStream_t)stream;
    uint64_t objects = 0;
    for (uint32_t k = 0; k < n_jobs; ++k) {
        Scene* s = ctx->scene(jobs[k].scene);
        if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
        objects += (jobs[k].object_begin == 0 && jobs[k].object_count == 0) ? s->n_objects : jobs[k].object_count;
    }
    uint32_t bucket = 0;
    while ((objects >> (bucket + 1)) != 0) ++bucket;
    const std::string key =
        std::string(ctx->device_name) + "|k" + std::to_string(kernel) + "|b" + std::to_string(bucket);
    std::map<std::string, CacheEntry> cache;
    if (cache_path) {
        cache = read_cache(cache_path);
        auto it = cache.find(key);
        if (it != cache.end() && valid_tuning(it->second.t)) {
            ctx->tune[kernel] = it->second.t;
            *best = it->second.t;
            return BZ_OK;
        }
    }

    auto run = [&]() -> bz_status {
        if (kernel == 1) return bz_cull_batch(ctx, jobs, n_jobs, stream);
        for (uint32_t k = 0; k < n_jobs; ++k) {
            const bz_view_job& j = jobs[k];
            bz_cull_args a;
            std::memset(&a, 0, sizeof a);
            a.scene = j.scene;
            a.pyramid = j.pyramid;
            a.pass = j.pass;
            a.view = &j.view;
            a.cmds_dev = j.cmds_dev;
            a.count_dev = j.count_dev;
            a.capacity = j.capacity;
            a.object_begin = j.object_begin;
            a.object_count = j.object_count;
            a.visibility_in_dev = j.visibility_in_dev;
            a.visibility_out_dev = j.visibility_out_dev;
            bz_status s = bz_cull(ctx, &a, stream);
            if (s) return s;
        }
        return BZ_OK;
    };

    std::vector<bz_tuning> cands;
    const uint32_t per_cu[] = { 0, 1, 2, 4 };
    for (uint32_t b : { 256u, 512u })
        for (uint32_t i : { 2u, 4u })
            for (uint32_t p : per_cu) {
                if (kernel == 1 && p) continue;  // the batched decide kernel launches one workgroup per tile
                cands.push_back(bz_tuning{ b, i, p, 0 });
            }

    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess) return ctx->hip_fail(he, "hipEventCreate");
    const bz_tuning saved = ctx->tune[kernel];
    bz_tuning winner = saved;
    double best_us = 1e300;
    bz_status status = BZ_OK;
    for (const bz_tuning& c : cands) {
        ctx->tune[kernel] = c;
        // warm-up (also grows the scratch), then a short timed run sized to about 2 ms
        float ms = 0.f;
        if ((status = run())) break;
        he = hipEventRecord(e0, st);
        if (he == hipSuccess && (status = run()) == BZ_OK) he = hipEventRecord(e1, st);
        if (status) break;
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess) break;
        const int reps = std::max(3, std::min(50, (int)std::ceil(2.0 / std::max(ms, 1e-3f))));
        he = hipEventRecord(e0, st);
        for (int r = 0; r < reps && he == hipSuccess && status == BZ_OK; ++r) status = run();
        if (status) break;
        if (he == hipSuccess) he = hipEventRecord(e1, st);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess) break;
        const double us = 1000.0 * ms / reps;
        if (us < best_us) {
            best_us = us;
            winner = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (status) {
        ctx->tune[kernel] = saved;
        return status;
    }
    if (he != hipSuccess) {
        ctx->tune[kernel] = saved;
        return ctx->hip_fail(he, "autotune timing");
    }
    ctx->tune[kernel] = winner;
    *best = winner;
    if (cache_path) {
        cache[key] = CacheEntry{ winner, best_us };
        if (!write_cache(cache_path, cache))
            return ctx->fail(BZ_E_INVALID, std::string("cannot write the cache file ") + cache_path);
    }
    return BZ_OK;
}
