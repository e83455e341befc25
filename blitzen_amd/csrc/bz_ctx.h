// bz_ctx.h — host-side state behind the C ABI: resident scenes, pyramids, per-stream scratch, graphs, server.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bz_internal.h"

namespace bz {

struct Scene {
    uint32_t n_objects = 0, n_transforms = 0, n_surfaces = 0, n_lods = 0, n_clusters = 0;
    uint32_t max_lod_clusters = 0;   // largest LodData.clusterCount (set with the cluster table)
    bool xf_identity = false;
    uint32_t* obj_surf = nullptr;      // surface ids, 32-bit (more than 65 536 surfaces) ...
    uint16_t* obj_surf16 = nullptr;    // ... or 16-bit (exactly one of the two is set)
    uint32_t* obj_xf = nullptr;
    float4* xf_ps = nullptr;
    float4* xf_q = nullptr;
    float4* surf_sphere = nullptr;
    uint2* surf_lod = nullptr;
    float* lod_error = nullptr;
    uint2* lod_draw = nullptr;
    uint2* lod_cluster = nullptr;
    uint32_t* vis = nullptr;
    float4* cl_sphere = nullptr;
    uint32_t* cl_cone = nullptr;
    uint2* cl_draw = nullptr;
    uint64_t bytes = 0;
    uint64_t xf_version = 0;            // bumped by every transform update (invalidates the frame cache)
    uint32_t* xf_winner = nullptr;      // [n_transforms] last-occurrence resolution of a commit's ids, zero between
    // frame cache: the frustum survivors of the last EARLY pass that recorded one (BZ_CULL_FRAME_CACHE); valid for
    // the same view, object range and transforms (the visibility is read by the LATE pass itself)
    struct FrameCache {
        float4* sph = nullptr;
        uint2* meta = nullptr;
        uint32_t* counts = nullptr;
        uint64_t cap_entries = 0;
        uint32_t cap_tiles = 0;
        bool valid = false;
        float view_key[19] = {};        // view matrix rows, frustum constants and lodTarget the survivors used
        uint32_t begin = 0, n = 0, block = 0, items = 0;
        uint64_t xf_version = 0;
        uint64_t recorded = 0, reused = 0;  // EARLY passes that recorded, LATE passes that read the records
    } cache;
    // transform-update staging (bz_scene_transform_staging, also behind bz_scene_update_transforms): two pinned slots
    // [ids | 32-byte MeshTransforms], used in turn
    struct XfSlot {
        uint8_t* host = nullptr;       // pinned, read by the scatter kernel: cap ids, then cap transforms at xf_offset
        uint32_t* dev_ids = nullptr;   // [cap] device copy of the committed ids
        uint32_t cap = 0;
        size_t xf_offset = 0;
        hipEvent_t done = nullptr;     // the last commit from this slot completed
    } xf_slot[2];
    uint32_t xf_next = 0;              // slot handed out by the next staging call
    int xf_open = -1;                  // slot handed out and not yet committed
    SceneDev dev() const;
    ~Scene();
};

// Level layout and single-pass tiling of a min-depth pyramid (no storage).
struct PyrPlan {
    uint32_t base_w = 0, base_h = 0, mips = 0;
    uint32_t level_offset[kMaxMips] = {};
    uint64_t n_texels = 0;
    uint32_t tile_w = 0, tile_h = 0, tiles_x = 0, tiles_y = 0, tile_levels = 0;
    bool single_pass = false;
    PyramidDesc desc(const float* texels) const;
    PyrBuildJob build_job(const float* depth, uint32_t w, uint32_t h, float* texels, unsigned long long* tail) const;
};
// Fills `out` for a base_w x base_h pyramid of `mips` levels; false if it holds more than 2^32 - 1 texels.
bool plan_pyramid(uint32_t base_w, uint32_t base_h, uint32_t mips, PyrPlan* out);

struct Pyramid : PyrPlan {
    float* texels = nullptr;
    unsigned long long* tail = nullptr;  // [tiles_x * tiles_y] flagged per-tile words of the single-pass build
    PyramidDesc desc() const { return PyrPlan::desc(texels); }
    PyrBuildJob build_job(const float* depth, uint32_t w, uint32_t h) const
    {
        return PyrPlan::build_job(depth, w, h, texels, tail);
    }
    ~Pyramid();
};

struct StreamScratch {
    CompactScratch cs{};
    uint32_t* err = nullptr;          // [0] instanced scan wait gave up, [1] pyramid tail wait gave up, [2] staged transform id out of range
    CullJob* jobs_dev = nullptr;
    CullJob* jobs_host = nullptr;     // pinned
    uint32_t jobs_cap = 0;
    PyrBuildJob* pjobs_dev = nullptr;
    PyrBuildJob* pjobs_host = nullptr;
    uint32_t pjobs_cap = 0;
    hipEvent_t upload_done = nullptr; // last job-table upload from the pinned buffers
    ClusterScratch xs{};              // cluster pass
    uint32_t* inst_hist = nullptr;    // LOD-instanced output: chunk x LOD histogram, then each chunk's first slots
    uint64_t inst_hist_cap = 0;
    unsigned long long* inst_seg = nullptr;  // LOD-instanced output: the scan's published segment totals (zero)
    uint64_t inst_seg_cap = 0;
    bool inst_seg_dirty = false;      // a launch sequence stopped between the scan and the scatter that clears them
    ~StreamScratch();
};

struct Graph {
    hipGraphExec_t exec = nullptr;
    CompactScratch cs{};
    uint32_t* err = nullptr;          // as StreamScratch::err, for the graph's own kernels
    hipStream_t last_stream = nullptr;  // the stream of the last launch (bz_check on it reads `err`)
    bool launched = false;
    CullJob* jobs_dev = nullptr;
    PyrBuildJob* pjobs_dev = nullptr;
    ~Graph();
};

struct Server;  // server.cpp

// Memory imported from another API (a Vulkan buffer exported as a file descriptor) and the device pointers mapped
// onto it. `fd` is this library's duplicate of the caller's descriptor while the runtime has not taken it over.
struct ExtMem {
    hipExternalMemory_t mem = nullptr;
    int fd = -1;
    uint64_t size = 0;
    std::vector<void*> maps;
    ~ExtMem();  // interop.cpp
};

struct ExtSem {
    hipExternalSemaphore_t sem = nullptr;
    int fd = -1;
    uint32_t kind = 0;
    ~ExtSem();  // interop.cpp
};

struct Ctx {
    int device = 0;
    uint32_t n_cu = 0;
    uint64_t hbm_bytes = 0;
    uint32_t wait_limit = BZ_DEFAULT_WAIT_LIMIT;  // sleeps before a kernel's bounded wait gives up
    char device_name[64] = {};
    std::string last_error;
    std::mutex err_mu;
    uint64_t next_handle = 0;
    std::unordered_map<uint64_t, std::unique_ptr<Scene>> scenes;
    std::unordered_map<uint64_t, std::unique_ptr<Pyramid>> pyramids;
    std::unordered_map<uint64_t, std::unique_ptr<Graph>> graphs;
    std::unordered_map<uint64_t, std::unique_ptr<ExtMem>> extmems;
    std::unordered_map<uint64_t, std::unique_ptr<ExtSem>> extsems;
    std::unordered_map<uintptr_t, std::unique_ptr<StreamScratch>> scratch;
    bz_tuning tune[2] = {};
    std::unique_ptr<Server> server;
    void* noalloc_scratch = nullptr;  // device address of this device's fixed scratch (bz_cull_noalloc)

    Ctx();
    ~Ctx();

    bz_status fail(bz_status s, const std::string& what);
    bz_status hip_fail(hipError_t e, const char* what);
    void free_later(void* p);
    Scene* scene(bz_scene h);
    Pyramid* pyramid(bz_pyramid h);
    bz_status scratch_for(hipStream_t st, uint32_t tiles, uint64_t entries, StreamScratch** out);
    bz_status cluster_scratch_for(hipStream_t st, StreamScratch* sc, uint32_t tiles, uint64_t records);
    bz_status make_job(uint32_t pass, uint32_t flags, bz_scene sh, bz_pyramid ph, const bz_CameraViewData& view,
                       uint32_t begin, uint32_t count, const uint32_t* vis_in, uint32_t* vis_out, void* cmds,
                       uint32_t* count_dev, uint32_t capacity, uint32_t block, uint32_t items, CullJob* job);
    // Jobs with objects go to `out` (tile_start set); the count buffers of jobs with none go to `empty`.
    bz_status make_batch(const bz_view_job* jobs, uint32_t n_jobs, const bz_tuning& t, std::vector<CullJob>& out,
                         std::vector<uint32_t*>& empty, uint32_t* total_tiles);
    uint32_t batch_grid(const bz_tuning& t, uint32_t total_tiles) const;
    bz_status upload_jobs(StreamScratch* sc, const CullJob* jobs, uint32_t n, hipStream_t st);
    bz_status upload_pjobs(StreamScratch* sc, const PyrBuildJob* jobs, uint32_t n, hipStream_t st);
    bz_status build_pyramid(Pyramid* p, const float* depth, uint32_t w, uint32_t h, hipStream_t st);
    // bz_cull; with `fixed`, that compaction scratch is used and nothing is allocated (bz_cull_noalloc)
    bz_status cull(const bz_cull_args* a, hipStream_t st, const CompactScratch* fixed);
};

ViewConsts make_view_consts(const bz_CameraViewData& v);
// Device address of the fixed scratch array on `device` (the current device), looked up once per device.
hipError_t noalloc_scratch_address(int device, void** out);

// Job tables live in one allocation: cap descriptors, then cap uint32 first-tile (or first-workgroup) indices, which
// the batched kernels stage in LDS for their job lookup. The host copy is pinned and reused once the previous
// upload from it has completed.
template <class T>
inline bz_status upload_table(Ctx* ctx, T*& dev, T*& host, uint32_t& cap, hipEvent_t done, const T* jobs,
                              const uint32_t* starts, uint32_t n, hipStream_t st, const char* what)
{
    (void)hipEventSynchronize(done);
    if (cap < n) {
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
        dev = nullptr;
        host = nullptr;
        cap = 0;
        const uint32_t c = std::max<uint32_t>(n, 256u);
        const size_t bytes = (size_t)c * (sizeof(T) + 4);
        hipError_t e = hipMalloc((void**)&dev, bytes);
        if (e == hipSuccess) e = hipHostMalloc((void**)&host, bytes, hipHostMallocDefault);
        if (e != hipSuccess) return ctx->hip_fail(e, what);
        cap = c;
    }
    char* hs = reinterpret_cast<char*>(host) + (size_t)cap * sizeof(T);
    char* ds = reinterpret_cast<char*>(dev) + (size_t)cap * sizeof(T);
    std::memcpy(host, jobs, (size_t)n * sizeof(T));
    std::memcpy(hs, starts, (size_t)n * 4);
    hipError_t e = hipMemcpyAsync(dev, host, (size_t)n * sizeof(T), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, hs, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return ctx->hip_fail(e, what);
    return BZ_OK;
}

template <class T>
inline const uint32_t* table_starts(const T* dev, uint32_t cap)
{
    return reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(dev) + (size_t)cap * sizeof(T));
}

uint32_t tiles_for(uint32_t n, uint32_t block, uint32_t items);

}  // namespace bz

struct bz_ctx : public bz::Ctx {};
