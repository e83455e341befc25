// capi.cpp — the C ABI of bzcull (declared in include/bzcull.h).
//
// Host-side replacement of the Vulkan renderer's cull plumbing: buffer residency (vulkanRendererSetup.cpp:313-363,
// 365-647), per-pass count reset + dispatch (vulkanDraw.cpp:107-226), depth pyramid creation and generation
// (vulkanResources.cpp:85-116, vulkanDraw.cpp:554-622) and the per-frame transform stream (vulkanDraw.cpp:46-60,
// 816-820). Every entry point validates its arguments, maps HIP failures to status codes and never throws.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bz_ctx.h"

using namespace bz;

static_assert(sizeof(bz_IndirectDrawCmd32) == 32, "IndirectDrawCmd is 32 bytes (dx12Data.h:284-296)");
static_assert(sizeof(bz_LodInstanceCounter) == 8, "LodInstanceCounter is 8 bytes");
static_assert(sizeof(bz_RenderObject) == 8, "RenderObject layout");
static_assert(sizeof(bz_MeshTransform) == 32, "MeshTransform layout");
static_assert(sizeof(bz_PrimitiveSurface) == 32, "PrimitiveSurface layout");
static_assert(sizeof(bz_LodData) == 32, "LodData layout");
static_assert(sizeof(bz_Cluster) == 32, "Cluster layout");
static_assert(sizeof(bz_IndirectDrawData) == 24, "IndirectDrawData layout");
static_assert(sizeof(bz_ClusterDispatchData) == 12, "ClusterDispatchData layout");
static_assert(sizeof(bz_CameraViewData) == 256, "CameraViewData layout");
static_assert(sizeof(bz_result) == 64 && offsetof(bz_result, arrival_ns) == 56, "bz_result layout");
static_assert(offsetof(bz_CameraViewData, lodTarget) == 180, "CameraViewData.lodTarget offset");
static_assert(offsetof(bz_CameraViewData, frustumRight) == 140, "CameraViewData.frustumRight offset");
static_assert(offsetof(bz_Cluster, dataOffset) == 20, "Cluster.dataOffset offset");

namespace bz {

bz_status Ctx::fail(bz_status s, const std::string& what)
{
    std::lock_guard<std::mutex> g(err_mu);
    last_error = what;
    return s;
}

bz_status Ctx::hip_fail(hipError_t e, const char* what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();  // clear the sticky launch error
    return fail(e == hipErrorOutOfMemory ? BZ_E_OOM : BZ_E_HIP, m);
}

void Ctx::free_later(void* p)
{
    if (p) (void)hipFree(p);
}

Scene* Ctx::scene(bz_scene h)
{
    auto it = scenes.find(h);
    return it == scenes.end() ? nullptr : it->second.get();
}

Pyramid* Ctx::pyramid(bz_pyramid h)
{
    auto it = pyramids.find(h);
    return it == pyramids.end() ? nullptr : it->second.get();
}

Scene::~Scene()
{
    void* ps[] = { obj_surf, obj_surf16, obj_xf, xf_ps, xf_q, surf_sphere, surf_lod, lod_error, lod_draw, lod_cluster, vis,
                   cl_sphere, cl_cone, cl_draw, cache.sph, cache.meta, cache.counts, xf_winner };
    for (XfSlot& x : xf_slot)
        if (x.done) (void)hipEventSynchronize(x.done);
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (XfSlot& x : xf_slot) {
        if (x.dev_ids) (void)hipFree(x.dev_ids);
        if (x.host) (void)hipHostFree(x.host);
        if (x.done) (void)hipEventDestroy(x.done);
    }
}

Pyramid::~Pyramid()
{
    if (texels) (void)hipFree(texels);
    if (tail) (void)hipFree(tail);
}

StreamScratch::~StreamScratch()
{
    void* ps[] = { cs.counts, cs.staging, jobs_dev, pjobs_dev, err, xs.ccounts, xs.rbase, xs.scount, xs.sprefix,
                   xs.survivors, inst_hist, inst_seg };
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (jobs_host) (void)hipHostFree(jobs_host);
    if (pjobs_host) (void)hipHostFree(pjobs_host);
    if (upload_done) (void)hipEventDestroy(upload_done);
}

// Compaction scratch for `stream`, sized for at least `tiles` tiles and `entries` staging entries. Growth
// synchronizes the stream first and is refused while the stream is being captured into a graph.
bz_status Ctx::scratch_for(hipStream_t st, uint32_t tiles, uint64_t entries, StreamScratch** out)
{
    auto& slot = scratch[(uintptr_t)st];
    if (!slot) slot.reset(new StreamScratch());
    StreamScratch& s = *slot;
    if (!s.err) {
        hipError_t e = hipMalloc((void**)&s.err, 16);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(error flags)");
        e = hipMemset(s.err, 0, 16);
        if (e != hipSuccess) return hip_fail(e, "hipMemset(error flags)");
        e = hipEventCreateWithFlags(&s.upload_done, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    }
    if (s.cs.capacity_tiles < tiles || s.cs.capacity_entries < entries) {
        hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cst) == hipSuccess && cst != hipStreamCaptureStatusNone)
            return fail(BZ_E_STATE, "compaction scratch must grow, which is not allowed during stream capture; "
                                    "run the call once before capturing");
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        const uint32_t cap_t = std::max<uint32_t>(std::max(tiles, s.cs.capacity_tiles), 1u << 12);
        const uint64_t cap_e = std::max<uint64_t>(std::max(entries, s.cs.capacity_entries), 1ull << 20);
        if (s.cs.counts) (void)hipFree(s.cs.counts);
        if (s.cs.staging) (void)hipFree(s.cs.staging);
        s.cs = CompactScratch{};
        e = hipMalloc((void**)&s.cs.counts, (size_t)counts_words(cap_t) * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.cs.staging, (size_t)cap_e * sizeof(uint2));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(compaction scratch)");
        s.cs.capacity_tiles = cap_t;
        s.cs.capacity_entries = cap_e;
        set_prefix(s.cs);
    }
    *out = &s;
    return BZ_OK;
}

// Cluster-pass scratch: per-tile arrays for `tiles` object tiles and a survivor buffer of `records` entries.
bz_status Ctx::cluster_scratch_for(hipStream_t st, StreamScratch* sc, uint32_t tiles, uint64_t records)
{
    ClusterScratch& x = sc->xs;
    if (x.capacity_tiles >= tiles && x.capacity_records >= records && x.ccounts) return BZ_OK;
    hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cst) == hipSuccess && cst != hipStreamCaptureStatusNone)
        return fail(BZ_E_STATE, "cluster scratch must grow, which is not allowed during stream capture");
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    const uint32_t cap_t = std::max<uint32_t>(std::max(tiles, x.capacity_tiles), 1u << 10);
    const uint64_t cap_r = std::max<uint64_t>(std::max(records, x.capacity_records), 1ull << 16);
    for (void* p : { (void*)x.ccounts, (void*)x.rbase, (void*)x.scount, (void*)x.sprefix, (void*)x.survivors })
        if (p) (void)hipFree(p);
    x = ClusterScratch{};
    e = hipMalloc((void**)&x.ccounts, (size_t)cap_t * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&x.rbase, ((size_t)cap_t + 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&x.scount, (size_t)cap_t * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&x.sprefix, ((size_t)cap_t + 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&x.survivors, (size_t)cap_r * sizeof(uint2));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(cluster scratch)");
    x.capacity_tiles = cap_t;
    x.capacity_records = cap_r;
    return BZ_OK;
}

ViewConsts make_view_consts(const bz_CameraViewData& v)
{
    ViewConsts c;
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 4; ++col) c.m[4 * r + col] = v.view[col * 4 + r];
    c.fR = v.frustumRight;
    c.fL = v.frustumLeft;
    c.fT = v.frustumTop;
    c.fB = v.frustumBottom;
    c.p00 = v.proj0;
    c.p11 = v.proj5;
    c.znear = v.zNear;
    c.zfar = v.zFar;
    c.pyrW = v.pyramidWidth;
    c.pyrH = v.pyramidHeight;
    c.lodTarget = v.lodTarget;
    c.pad = 0.0f;
    return c;
}

SceneDev Scene::dev() const
{
    SceneDev d;
    d.obj_surf = obj_surf;
    d.obj_surf16 = obj_surf16;
    d.obj_xf = obj_xf;
    d.xf_ps = xf_ps;
    d.xf_q = xf_q;
    d.surf_sphere = surf_sphere;
    d.surf_lod = surf_lod;
    d.lod_error = lod_error;
    d.lod_draw = lod_draw;
    d.lod_cluster = lod_cluster;
    d.n_objects = n_objects;
    d.n_transforms = n_transforms;
    d.n_surfaces = n_surfaces;
    d.n_lods = n_lods;
    return d;
}

PyramidDesc PyrPlan::desc(const float* texels) const
{
    PyramidDesc p;
    std::memset(&p, 0, sizeof p);
    p.texels = texels;
    p.base_w = base_w;
    p.base_h = base_h;
    p.mips = mips;
    for (uint32_t i = 0; i < mips && i < (uint32_t)kMaxMips; ++i) p.level_offset[i] = level_offset[i];
    return p;
}

bool plan_pyramid(uint32_t base_w, uint32_t base_h, uint32_t mips, PyrPlan* p)
{
    *p = PyrPlan{};
    p->base_w = base_w;
    p->base_h = base_h;
    p->mips = mips;
    uint64_t off = 0;
    for (uint32_t L = 0; L < mips; ++L) {
        p->level_offset[L] = (uint32_t)off;
        off += (uint64_t)level_dim(base_w, L) * level_dim(base_h, L);
    }
    if (off > 0xffffffffull) return false;
    p->n_texels = off;
    const bool pow2 = (base_w & (base_w - 1)) == 0 && (base_h & (base_h - 1)) == 0;
    p->tile_w = std::min<uint32_t>(32u, base_w);
    p->tile_h = std::min<uint32_t>(32u, base_h);
    p->tiles_x = base_w / p->tile_w;
    p->tiles_y = base_h / p->tile_h;
    uint32_t tl = 1, m = std::max(p->tile_w, p->tile_h);
    while (m > 1) {
        m >>= 1;
        ++tl;
    }
    p->tile_levels = std::min(tl, mips);
    // single pass: power-of-two extent, and the tile image's first reduction fits the tail workgroup's LDS tile
    const uint64_t tail_texels = (uint64_t)std::max(1u, p->tiles_x / 2u) * std::max(1u, p->tiles_y / 2u);
    p->single_pass = pow2 && (p->tile_levels >= mips || tail_texels <= 1024u);
    return true;
}

uint32_t tiles_for(uint32_t n, uint32_t block, uint32_t items)
{
    const uint64_t per = (uint64_t)block * items;
    const uint64_t t = (n + per - 1) / per;
    return t ? (uint32_t)t : 1u;
}

// Validates one view's cull and fills a CullJob (without tile_start).
bz_status Ctx::make_job(uint32_t pass, uint32_t flags, bz_scene sh, bz_pyramid ph, const bz_CameraViewData& view,
                        uint32_t begin, uint32_t count, const uint32_t* vis_in, uint32_t* vis_out, void* cmds,
                        uint32_t* count_dev, uint32_t capacity, uint32_t block, uint32_t items, CullJob* job)
{
    Scene* s = scene(sh);
    if (!s) return fail(BZ_E_INVALID, "unknown scene handle");
    if (pass > PASS_ONPC) return fail(BZ_E_INVALID, "unknown pass");
    if (flags & ~(BZ_CULL_EARLY_HIZ | BZ_CULL_FRAME_CACHE)) return fail(BZ_E_INVALID, "unknown cull flags");
    if (!count_dev) return fail(BZ_E_INVALID, "count_dev is NULL");
    if (capacity && !cmds) return fail(BZ_E_INVALID, "cmds_dev is NULL with capacity > 0");
    if ((uintptr_t)cmds % 8u) return fail(BZ_E_INVALID, "cmds_dev must be 8-byte aligned");
    if (begin == 0 && count == 0) count = s->n_objects;
    if ((uint64_t)begin + count > s->n_objects) return fail(BZ_E_INVALID, "object range exceeds the scene");
    Pyramid* p = nullptr;
    if (ph) {
        p = pyramid(ph);
        if (!p) return fail(BZ_E_INVALID, "unknown pyramid handle");
    }
    const bool hiz = (pass == PASS_LATE || pass == PASS_TEMPORAL || (pass == PASS_EARLY && (flags & BZ_CULL_EARLY_HIZ)));
    if (pass == PASS_LATE && !p) return fail(BZ_E_INVALID, "the LATE pass needs a pyramid");
    if (pass == PASS_EARLY && (flags & BZ_CULL_EARLY_HIZ) && !p)
        return fail(BZ_E_INVALID, "BZ_CULL_EARLY_HIZ needs a pyramid");
    std::memset(job, 0, sizeof *job);
    job->scene = s->dev();
    job->view = make_view_consts(view);
    if (hiz && p) job->pyr = p->desc();
    job->hiz = (hiz && p) ? 1u : 0u;
    job->vis_in = vis_in ? vis_in : s->vis;
    job->vis_out = vis_out ? vis_out : const_cast<uint32_t*>(job->vis_in);
    job->cmds = reinterpret_cast<uint32_t*>(cmds);
    job->count = count_dev;
    job->begin = begin;
    job->n = count;
    job->capacity = capacity;
    job->pass = pass;
    job->n_tiles = tiles_for(count, block, items);
    return BZ_OK;
}

}  // namespace bz

// ================================================================================================
// C ABI
// ================================================================================================
extern "C" {

uint32_t bz_abi_version(void) { return BZCULL_ABI_VERSION; }

const char* bz_status_str(bz_status s)
{
    switch (s) {
    case BZ_OK: return "ok";
    case BZ_E_INVALID: return "invalid argument";
    case BZ_E_OOM: return "out of memory";
    case BZ_E_HIP: return "HIP runtime error";
    case BZ_E_CAPACITY: return "output capacity exceeded";
    case BZ_E_BUDGET: return "HBM budget exceeded";
    case BZ_E_TIMEOUT: return "timeout";
    case BZ_E_NOT_READY: return "not ready";
    case BZ_E_NOT_FOUND: return "not found";
    case BZ_E_STATE: return "invalid state";
    case BZ_E_NODEVICE: return "no device";
    case BZ_E_INTERNAL: return "internal kernel error";
    case BZ_E_CANCELLED: return "cancelled";
    default: return "unknown status";
    }
}

const char* bz_last_error(const bz_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

bz_status bz_create(const bz_config* cfg, bz_ctx** out)
{
    if (!out) return BZ_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return BZ_E_NODEVICE;
    }
    const int dev = cfg ? cfg->device : 0;
    if (dev < 0 || dev >= ndev) return BZ_E_INVALID;
    if (hipSetDevice(dev) != hipSuccess) return BZ_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return BZ_E_HIP;
    std::unique_ptr<bz_ctx> c(new (std::nothrow) bz_ctx());
    if (!c) return BZ_E_OOM;
    c->device = dev;
    if (noalloc_scratch_address(dev, &c->noalloc_scratch) != hipSuccess) {  // bz_cull_noalloc's fixed scratch
        (void)hipGetLastError();
        return BZ_E_HIP;
    }
    c->n_cu = (uint32_t)prop.multiProcessorCount;
    c->hbm_bytes = (uint64_t)prop.totalGlobalMem;
    std::snprintf(c->device_name, sizeof c->device_name, "%s", prop.gcnArchName);
    c->tune[0] = bz_tuning{ 512, 2, 0, 0 };
    c->tune[1] = bz_tuning{ 256, 4, 8, 0 };
    *out = c.release();
    return BZ_OK;
}

void bz_destroy(bz_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    bz_server_stop(ctx);
    (void)hipDeviceSynchronize();
    delete ctx;
}

bz_status bz_device_info(bz_ctx* ctx, char* name, uint32_t name_len, uint32_t* n_cu, uint64_t* hbm_bytes)
{
    if (!ctx) return BZ_E_INVALID;
    if (name && name_len) std::snprintf(name, name_len, "%s", ctx->device_name);
    if (n_cu) *n_cu = ctx->n_cu;
    if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
    return BZ_OK;
}

// ---------------------------------------------------------------------------- scenes
bz_status bz_scene_create(bz_ctx* ctx, const bz_RenderObject* objects, uint32_t n_objects,
                          const bz_MeshTransform* transforms, uint32_t n_transforms,
                          const bz_PrimitiveSurface* surfaces, uint32_t n_surfaces, const bz_LodData* lods,
                          uint32_t n_lods, bz_scene* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    *out = 0;
    if (n_objects > BZ_MAX_OBJECTS)
        return ctx->fail(BZ_E_INVALID, "more than BZ_MAX_OBJECTS objects in one scene");
    if ((n_objects && !objects) || (n_transforms && !transforms) || (n_surfaces && !surfaces) || (n_lods && !lods))
        return ctx->fail(BZ_E_INVALID, "NULL array with non-zero count");
    if (n_surfaces == 0 || n_lods == 0 || n_transforms == 0)
        return ctx->fail(BZ_E_INVALID, "a scene needs at least one transform, surface and LOD");
    // Validation: the kernels never bounds-check ids, so every id is checked here once.
    for (uint32_t s = 0; s < n_surfaces; ++s) {
        const bz_PrimitiveSurface& sf = surfaces[s];
        if (sf.lodOffset >= n_lods || (uint64_t)sf.lodOffset + sf.lodCount > n_lods)
            return ctx->fail(BZ_E_INVALID, "surface " + std::to_string(s) + " LOD range exceeds the LOD table");
    }
    bool identity = n_transforms >= n_objects;
    for (uint32_t i = 0; i < n_objects; ++i) {
        if (objects[i].transformId >= n_transforms)
            return ctx->fail(BZ_E_INVALID, "object " + std::to_string(i) + " transformId out of range");
        if (objects[i].surfaceId >= n_surfaces)
            return ctx->fail(BZ_E_INVALID, "object " + std::to_string(i) + " surfaceId out of range");
        identity = identity && objects[i].transformId == i;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(BZ_E_HIP, "hipSetDevice");
    std::unique_ptr<Scene> s(new (std::nothrow) Scene());
    if (!s) return BZ_E_OOM;
    s->n_objects = n_objects;
    s->n_transforms = n_transforms;
    s->n_surfaces = n_surfaces;
    s->n_lods = n_lods;
    s->xf_identity = identity;
    // host SoA staging
    // surface ids as 16-bit when every id fits (any table of at most 65 536 surfaces): 2 B instead of 4 B per object
    // on every pass over the scene
    const bool surf16 = n_surfaces <= 65536u;
    std::vector<uint32_t> surf_ids(surf16 ? 0 : n_objects), xf_ids(identity ? 0 : n_objects);
    std::vector<uint16_t> surf_ids16(surf16 ? n_objects : 0);
    for (uint32_t i = 0; i < n_objects; ++i) {
        if (surf16) surf_ids16[i] = (uint16_t)objects[i].surfaceId;
        else surf_ids[i] = objects[i].surfaceId;
        if (!identity) xf_ids[i] = objects[i].transformId;
    }
    std::vector<float4> ps(n_transforms), q(n_transforms), sph(n_surfaces);
    std::vector<uint2> slod(n_surfaces), ldraw(n_lods), lcl(n_lods);
    std::vector<float> lerr(n_lods);
    for (uint32_t t = 0; t < n_transforms; ++t) {
        const bz_MeshTransform& m = transforms[t];
        ps[t] = make_float4(m.pos[0], m.pos[1], m.pos[2], m.scale);
        q[t] = make_float4(m.orientation[0], m.orientation[1], m.orientation[2], m.orientation[3]);
    }
    for (uint32_t k = 0; k < n_surfaces; ++k) {
        const bz_PrimitiveSurface& sf = surfaces[k];
        sph[k] = make_float4(sf.center[0], sf.center[1], sf.center[2], sf.radius);
        slod[k] = make_uint2(sf.lodOffset, sf.lodCount);
    }
    for (uint32_t l = 0; l < n_lods; ++l) {
        lerr[l] = lods[l].error;
        ldraw[l] = make_uint2(lods[l].indexCount, lods[l].firstIndex);
        lcl[l] = make_uint2(lods[l].clusterOffset, lods[l].clusterCount);
    }
    auto up = [&](void** dst, const void* src, size_t bytes) -> bz_status {
        if (bytes == 0) return BZ_OK;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return ctx->hip_fail(e, "hipMalloc(scene)");
        s->bytes += bytes;
        if (src) {
            e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) return ctx->hip_fail(e, "hipMemcpy(scene)");
        }
        return BZ_OK;
    };
    bz_status st;
    if (surf16) {
        if ((st = up((void**)&s->obj_surf16, surf_ids16.data(), surf_ids16.size() * 2))) return st;
    } else if ((st = up((void**)&s->obj_surf, surf_ids.data(), surf_ids.size() * 4))) {
        return st;
    }
    if (!identity && (st = up((void**)&s->obj_xf, xf_ids.data(), xf_ids.size() * 4))) return st;
    if ((st = up((void**)&s->xf_ps, ps.data(), ps.size() * sizeof(float4)))) return st;
    if ((st = up((void**)&s->xf_q, q.data(), q.size() * sizeof(float4)))) return st;
    if ((st = up((void**)&s->surf_sphere, sph.data(), sph.size() * sizeof(float4)))) return st;
    if ((st = up((void**)&s->surf_lod, slod.data(), slod.size() * sizeof(uint2)))) return st;
    if ((st = up((void**)&s->lod_error, lerr.data(), lerr.size() * 4))) return st;
    if ((st = up((void**)&s->lod_draw, ldraw.data(), ldraw.size() * sizeof(uint2)))) return st;
    if ((st = up((void**)&s->lod_cluster, lcl.data(), lcl.size() * sizeof(uint2)))) return st;
    if ((st = up((void**)&s->vis, nullptr, (size_t)std::max<uint32_t>(n_objects, 1u) * 4))) return st;
    // visibility starts at zero (vulkanRendererSetup.cpp:349)
    hipError_t e = launch_fill_u32(s->vis, 0u, std::max<uint32_t>(n_objects, 1u), nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return ctx->hip_fail(e, "visibility fill");
    const bz_scene h = ++ctx->next_handle;
    ctx->scenes[h] = std::move(s);
    *out = h;
    return BZ_OK;
}

bz_status bz_scene_set_clusters(bz_ctx* ctx, bz_scene scene, const bz_Cluster* clusters, uint32_t n_clusters)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (n_clusters && !clusters) return ctx->fail(BZ_E_INVALID, "NULL clusters");
    // every LOD's cluster range must be inside the table
    std::vector<uint2> lcl(s->n_lods);
    if (hipMemcpy(lcl.data(), s->lod_cluster, lcl.size() * sizeof(uint2), hipMemcpyDeviceToHost) != hipSuccess)
        return ctx->fail(BZ_E_HIP, "hipMemcpy(lod clusters)");
    uint32_t max_cl = 0;
    for (uint32_t l = 0; l < s->n_lods; ++l) {
        if (lcl[l].y && (uint64_t)lcl[l].x + lcl[l].y > n_clusters)
            return ctx->fail(BZ_E_INVALID, "LOD " + std::to_string(l) + " cluster range exceeds the cluster table");
        max_cl = std::max(max_cl, lcl[l].y);
    }
    std::vector<float4> sph(n_clusters);
    std::vector<uint32_t> cone(n_clusters);
    std::vector<uint2> draw(n_clusters);
    for (uint32_t c = 0; c < n_clusters; ++c) {
        const bz_Cluster& k = clusters[c];
        sph[c] = make_float4(k.center[0], k.center[1], k.center[2], k.radius);
        cone[c] = (uint32_t)(uint8_t)k.coneAxisX | ((uint32_t)(uint8_t)k.coneAxisY << 8) |
                  ((uint32_t)(uint8_t)k.coneAxisZ << 16) | ((uint32_t)(uint8_t)k.coneCutoff << 24);
        draw[c] = make_uint2((uint32_t)k.triangleCount * 3u, k.dataOffset);
    }
    for (void* p : { (void*)s->cl_sphere, (void*)s->cl_cone, (void*)s->cl_draw })
        if (p) (void)hipFree(p);
    s->cl_sphere = nullptr;
    s->cl_cone = nullptr;
    s->cl_draw = nullptr;
    s->n_clusters = n_clusters;
    s->max_lod_clusters = n_clusters ? max_cl : 0;
    if (!n_clusters) return BZ_OK;
    hipError_t e = hipMalloc((void**)&s->cl_sphere, sph.size() * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc((void**)&s->cl_cone, cone.size() * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->cl_draw, draw.size() * sizeof(uint2));
    if (e == hipSuccess) e = hipMemcpy(s->cl_sphere, sph.data(), sph.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->cl_cone, cone.data(), cone.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->cl_draw, draw.data(), draw.size() * sizeof(uint2), hipMemcpyHostToDevice);
    if (e != hipSuccess) return ctx->hip_fail(e, "cluster upload");
    return BZ_OK;
}

bz_status bz_scene_destroy(bz_ctx* ctx, bz_scene scene)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->scenes.find(scene);
    if (it == ctx->scenes.end()) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (ctx->server) return ctx->fail(BZ_E_STATE, "scenes cannot be destroyed while the server runs");
    (void)hipDeviceSynchronize();
    ctx->scenes.erase(it);
    return BZ_OK;
}

bz_status bz_scene_info(bz_ctx* ctx, bz_scene scene, uint32_t* n_objects, uint32_t* n_transforms, uint32_t* n_surfaces,
                        uint32_t* n_lods, uint64_t* device_bytes)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (n_objects) *n_objects = s->n_objects;
    if (n_transforms) *n_transforms = s->n_transforms;
    if (n_surfaces) *n_surfaces = s->n_surfaces;
    if (n_lods) *n_lods = s->n_lods;
    if (device_bytes) *device_bytes = s->bytes;
    return BZ_OK;
}

bz_status bz_scene_transform_staging(bz_ctx* ctx, bz_scene scene, uint32_t capacity, uint32_t** ids,
                                     bz_MeshTransform** transforms)
{
    if (!ctx || !ids || !transforms) return BZ_E_INVALID;
    *ids = nullptr;
    *transforms = nullptr;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (!capacity) return ctx->fail(BZ_E_INVALID, "capacity must be > 0");
    const uint32_t k = s->xf_open >= 0 ? (uint32_t)s->xf_open : s->xf_next;  // an open slot is handed out again
    Scene::XfSlot& x = s->xf_slot[k];
    if (x.done) {
        hipError_t e = hipEventSynchronize(x.done);  // the slot's previous upload has been read
        if (e != hipSuccess) return ctx->hip_fail(e, "staging slot wait");
    }
    if (x.cap < capacity) {
        if (x.host) (void)hipHostFree(x.host);
        if (x.dev_ids) (void)hipFree(x.dev_ids);
        x.host = nullptr;
        x.dev_ids = nullptr;
        x.cap = 0;
        const size_t off = ((size_t)capacity * 4 + 255) / 256 * 256;
        const size_t bytes = off + (size_t)capacity * sizeof(bz_MeshTransform);
        hipError_t e = hipHostMalloc((void**)&x.host, bytes, hipHostMallocDefault);  // pinned, device-readable
        if (e == hipSuccess) e = hipMalloc((void**)&x.dev_ids, (size_t)capacity * 4);
        if (e == hipSuccess && !s->xf_winner) {
            e = hipMalloc((void**)&s->xf_winner, (size_t)s->n_transforms * 4);
            if (e == hipSuccess) e = hipMemset(s->xf_winner, 0, (size_t)s->n_transforms * 4);
            if (e != hipSuccess && s->xf_winner) {
                (void)hipFree(s->xf_winner);
                s->xf_winner = nullptr;
            }
        }
        if (e == hipSuccess && !x.done) e = hipEventCreateWithFlags(&x.done, hipEventDisableTiming);
        if (e != hipSuccess) return ctx->hip_fail(e, "transform staging slot");
        x.cap = capacity;
        x.xf_offset = off;
    }
    s->xf_open = (int)k;
    *ids = reinterpret_cast<uint32_t*>(x.host);
    *transforms = reinterpret_cast<bz_MeshTransform*>(x.host + x.xf_offset);
    return BZ_OK;
}

bz_status bz_scene_commit_transforms(bz_ctx* ctx, bz_scene scene, uint32_t n, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (s->xf_open < 0) return ctx->fail(BZ_E_STATE, "no staging slot handed out (bz_scene_transform_staging)");
    Scene::XfSlot& x = s->xf_slot[s->xf_open];
    if (n > x.cap) return ctx->fail(BZ_E_INVALID, "more updates than the staging slot holds");
    hipStream_t st = (hipStream_t)stream;
    StreamScratch* sc = nullptr;
    bz_status status = ctx->scratch_for(st, 0, 0, &sc);  // the stream's error flags (no compaction scratch)
    if (status) return status;
    ++s->xf_version;  // frame-cache survivors recorded before this update no longer apply
    hipError_t e = hipSuccess;
    if (n) {
        // the kernels read the pinned slot directly over the host link (ids once, records once; no staging copy)
        e = launch_scatter_staged_transforms(reinterpret_cast<const uint32_t*>(x.host),
                                             reinterpret_cast<const float4*>(x.host + x.xf_offset), n,
                                             s->n_transforms, x.dev_ids, s->xf_winner, s->xf_ps, s->xf_q, sc->err,
                                             st);
    }
    if (e == hipSuccess) e = hipEventRecord(x.done, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "staged transform update");
    s->xf_next = (uint32_t)s->xf_open ^ 1u;
    s->xf_open = -1;
    return BZ_OK;
}

bz_status bz_scene_update_transforms(bz_ctx* ctx, bz_scene scene, const uint32_t* ids, const bz_MeshTransform* xf,
                                     uint32_t n, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (n == 0) return BZ_OK;
    if (!ids || !xf) return ctx->fail(BZ_E_INVALID, "NULL ids / transforms");
    for (uint32_t k = 0; k < n; ++k)
        if (ids[k] >= s->n_transforms) return ctx->fail(BZ_E_INVALID, "transform id out of range");
    if (s->xf_open >= 0) return ctx->fail(BZ_E_STATE, "a transform staging slot is open: commit it first");
    // through the staging slots: the host arrays are copied as they are (32-byte records), the kernel splits them
    uint32_t* sid = nullptr;
    bz_MeshTransform* sxf = nullptr;
    bz_status st = bz_scene_transform_staging(ctx, scene, n, &sid, &sxf);
    if (st) return st;
    std::memcpy(sid, ids, (size_t)n * 4);
    std::memcpy(sxf, xf, (size_t)n * sizeof(bz_MeshTransform));
    return bz_scene_commit_transforms(ctx, scene, n, stream);
}

bz_status bz_scene_set_visibility(bz_ctx* ctx, bz_scene scene, const uint32_t* vis, uint32_t n, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (n != s->n_objects || (n && !vis)) return ctx->fail(BZ_E_INVALID, "visibility size must equal n_objects");
    if (!n) return BZ_OK;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(s->vis, vis, (size_t)n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return ctx->hip_fail(e, "visibility upload");
    return BZ_OK;
}

bz_status bz_scene_get_visibility(bz_ctx* ctx, bz_scene scene, uint32_t* vis, uint32_t n, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (n != s->n_objects || (n && !vis)) return ctx->fail(BZ_E_INVALID, "visibility size must equal n_objects");
    if (!n) return BZ_OK;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(vis, s->vis, (size_t)n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return ctx->hip_fail(e, "visibility download");
    return BZ_OK;
}

bz_status bz_scene_get_transforms(bz_ctx* ctx, bz_scene scene, uint32_t first, uint32_t n, bz_MeshTransform* out,
                                  void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if ((uint64_t)first + n > s->n_transforms || (n && !out)) return ctx->fail(BZ_E_INVALID, "transform range");
    if (!n) return BZ_OK;
    std::vector<float4> ps(n), q(n);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(ps.data(), s->xf_ps + first, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(q.data(), s->xf_q + first, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return ctx->hip_fail(e, "transform download");
    for (uint32_t k = 0; k < n; ++k) {
        out[k].pos[0] = ps[k].x;
        out[k].pos[1] = ps[k].y;
        out[k].pos[2] = ps[k].z;
        out[k].scale = ps[k].w;
        out[k].orientation[0] = q[k].x;
        out[k].orientation[1] = q[k].y;
        out[k].orientation[2] = q[k].z;
        out[k].orientation[3] = q[k].w;
    }
    return BZ_OK;
}

bz_status bz_scene_frame_cache_stats(bz_ctx* ctx, bz_scene scene, uint64_t* recorded, uint64_t* reused)
{
    if (!ctx) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (recorded) *recorded = s->cache.recorded;
    if (reused) *reused = s->cache.reused;
    return BZ_OK;
}

bz_status bz_scene_visibility_ptr(bz_ctx* ctx, bz_scene scene, uint32_t** vis_dev)
{
    if (!ctx || !vis_dev) return BZ_E_INVALID;
    Scene* s = ctx->scene(scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    *vis_dev = s->vis;
    return BZ_OK;
}

// ---------------------------------------------------------------------------- pyramids
void bz_pyramid_reference_extent(uint32_t w, uint32_t h, uint32_t* base_w, uint32_t* base_h, uint32_t* mips)
{
    auto prev_pow2 = [](uint32_t v) {
        uint32_t r = 1;
        while (r * 2 < v) r *= 2;
        return r;
    };
    const uint32_t bw = prev_pow2(w), bh = prev_pow2(h);
    uint32_t n = 0, x = bw, y = bh;
    while (x > 1 || y > 1) {
        ++n;
        x /= 2;
        y /= 2;
    }
    if (base_w) *base_w = bw;
    if (base_h) *base_h = bh;
    if (mips) *mips = n;
}

bz_status bz_pyramid_create(bz_ctx* ctx, uint32_t base_w, uint32_t base_h, uint32_t mips, bz_pyramid* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    *out = 0;
    if (base_w == 0 || base_h == 0 || mips == 0 || mips > (uint32_t)kMaxMips)
        return ctx->fail(BZ_E_INVALID, "pyramid extent must be non-zero with 1..16 mips");
    std::unique_ptr<Pyramid> p(new (std::nothrow) Pyramid());
    if (!p) return BZ_E_OOM;
    if (!plan_pyramid(base_w, base_h, mips, p.get())) return ctx->fail(BZ_E_INVALID, "pyramid too large");
    const uint64_t off = p->n_texels;
    // one spare texel past the last level: the Hi-Z lookup reads texels in pairs (footprint_texels)
    hipError_t e = hipMalloc((void**)&p->texels, (off + 1) * 4);
    if (e == hipSuccess) e = hipMemset(p->texels, 0, (off + 1) * 4);
    const size_t n_tiles = (size_t)p->tiles_x * p->tiles_y;
    if (e == hipSuccess) e = hipMalloc((void**)&p->tail, n_tiles * 8);
    if (e == hipSuccess) e = hipMemset(p->tail, 0, n_tiles * 8);
    if (e != hipSuccess) return ctx->hip_fail(e, "pyramid allocation");
    const bz_pyramid h = ++ctx->next_handle;
    ctx->pyramids[h] = std::move(p);
    *out = h;
    return BZ_OK;
}

bz_status bz_pyramid_destroy(bz_ctx* ctx, bz_pyramid pyramid)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->pyramids.find(pyramid);
    if (it == ctx->pyramids.end()) return ctx->fail(BZ_E_INVALID, "unknown pyramid handle");
    (void)hipDeviceSynchronize();
    ctx->pyramids.erase(it);
    return BZ_OK;
}

bz_status bz_pyramid_storage(bz_ctx* ctx, bz_pyramid pyramid, float** texels, uint64_t* n_texels, uint32_t* offsets)
{
    if (!ctx) return BZ_E_INVALID;
    Pyramid* p = ctx->pyramid(pyramid);
    if (!p) return ctx->fail(BZ_E_INVALID, "unknown pyramid handle");
    if (texels) *texels = p->texels;
    if (n_texels) *n_texels = p->n_texels;
    if (offsets)
        for (uint32_t L = 0; L < p->mips; ++L) offsets[L] = p->level_offset[L];
    return BZ_OK;
}

namespace bz {
PyrBuildJob PyrPlan::build_job(const float* depth, uint32_t w, uint32_t h, float* texels,
                               unsigned long long* tail) const
{
    PyrBuildJob j;
    std::memset(&j, 0, sizeof j);
    j.depth = depth;
    j.w = w;
    j.h = h;
    j.texels = texels;
    j.base_w = base_w;
    j.base_h = base_h;
    j.mips = mips;
    j.tile_w = tile_w;
    j.tile_h = tile_h;
    j.tiles_x = tiles_x;
    j.tiles_y = tiles_y;
    j.tile_levels = tile_levels;
    j.tail = tail;
    j.spin_limit = BZ_DEFAULT_WAIT_LIMIT;
    for (uint32_t L = 0; L < mips; ++L) j.level_offset[L] = level_offset[L];
    return j;
}

bz_status Ctx::build_pyramid(Pyramid* p, const float* depth, uint32_t w, uint32_t h, hipStream_t st)
{
    StreamScratch* sc = nullptr;
    bz_status s = scratch_for(st, 0, 0, &sc);
    if (s) return s;
    hipError_t e = hipSuccess;
    if (p->single_pass) {
        PyrBuildJob j = p->build_job(depth, w, h);
        j.spin_limit = wait_limit;
        e = launch_pyramid(j, sc->err + 1, st);
    } else {
        e = launch_pyramid_level(depth, w, h, p->texels, p->base_w, p->base_h, st);
        for (uint32_t L = 1; L < p->mips && e == hipSuccess; ++L)
            e = launch_pyramid_level(p->texels + p->level_offset[L - 1], level_dim(p->base_w, L - 1),
                                     level_dim(p->base_h, L - 1), p->texels + p->level_offset[L],
                                     level_dim(p->base_w, L), level_dim(p->base_h, L), st);
    }
    if (e != hipSuccess) return hip_fail(e, "pyramid launch");
    return BZ_OK;
}
}  // namespace bz

bz_status bz_pyramid_build(bz_ctx* ctx, bz_pyramid pyramid, const float* depth, uint32_t w, uint32_t h, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    Pyramid* p = ctx->pyramid(pyramid);
    if (!p) return ctx->fail(BZ_E_INVALID, "unknown pyramid handle");
    if (!depth || w == 0 || h == 0) return ctx->fail(BZ_E_INVALID, "empty depth buffer");
    return ctx->build_pyramid(p, depth, w, h, (hipStream_t)stream);
}

bz_status bz_pyramid_build_batch(bz_ctx* ctx, const bz_pyramid_job* jobs, uint32_t n, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    if (n == 0) return BZ_OK;
    if (!jobs) return ctx->fail(BZ_E_INVALID, "NULL jobs");
    hipStream_t st = (hipStream_t)stream;
    std::vector<PyrBuildJob> sp;
    uint32_t wg = 0;
    for (uint32_t k = 0; k < n; ++k) {
        Pyramid* p = ctx->pyramid(jobs[k].pyramid);
        if (!p) return ctx->fail(BZ_E_INVALID, "unknown pyramid handle");
        if (!jobs[k].depth_dev || !jobs[k].w || !jobs[k].h) return ctx->fail(BZ_E_INVALID, "empty depth buffer");
        if (!p->single_pass) {
            bz_status s = ctx->build_pyramid(p, jobs[k].depth_dev, jobs[k].w, jobs[k].h, st);
            if (s) return s;
            continue;
        }
        PyrBuildJob j = p->build_job(jobs[k].depth_dev, jobs[k].w, jobs[k].h);
        j.spin_limit = ctx->wait_limit;
        j.wg_start = wg;
        wg += j.tiles_x * j.tiles_y;
        sp.push_back(j);
    }
    if (sp.empty()) return BZ_OK;
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return ctx->fail(BZ_E_STATE, "bz_pyramid_build_batch cannot be stream-captured; use bz_batch_graph_create");
    }
    StreamScratch* sc = nullptr;
    bz_status s = ctx->scratch_for(st, 0, 0, &sc);
    if (s) return s;
    s = ctx->upload_pjobs(sc, sp.data(), (uint32_t)sp.size(), st);
    if (s) return s;
    uint32_t mt = 0;
    for (auto& q : sp) mt = std::max<uint32_t>(mt, q.tiles_x * q.tiles_y);
    hipError_t e = launch_pyramid_batch(sc->pjobs_dev, (uint32_t)sp.size(), mt, sc->err + 1, st);
    if (e == hipSuccess) e = hipEventRecord(sc->upload_done, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "pyramid batch launch");
    return BZ_OK;
}

// ---------------------------------------------------------------------------- cull
namespace bz {
// One bz_cull call. With `fixed` the compaction scratch is that buffer (never grown) and nothing is allocated: a
// range that needs more, or a frame cache not yet allocated large enough, is refused.
bz_status Ctx::cull(const bz_cull_args* a, hipStream_t st, const CompactScratch* fixed)
{
    const bz_tuning& t = tune[0];
    CullJob job;
    bz_status s = make_job(a->pass, a->flags, a->scene, a->pyramid, *a->view, a->object_begin, a->object_count,
                           a->visibility_in_dev, a->visibility_out_dev, a->cmds_dev, a->count_dev, a->capacity, t.block,
                           t.items, &job);
    if (s) return s;
    if (job.n == 0) {  // nothing to cull: the count is 0 and the visibility buffers are untouched
        hipError_t e = launch_fill_u32(job.count, 0u, 1, st);
        if (e != hipSuccess) return hip_fail(e, "count fill");
        return BZ_OK;
    }
    // compaction scratch for `tiles` tiles of `entries` staging slots in all
    auto compact = [&](uint32_t tiles, uint64_t entries, CompactScratch* out) -> bz_status {
        if (fixed) {
            if (tiles > fixed->capacity_tiles || entries > fixed->capacity_entries)
                return fail(BZ_E_INVALID, "the range needs more compaction scratch than the fixed buffer holds");
            *out = *fixed;
            return BZ_OK;
        }
        StreamScratch* sc = nullptr;
        bz_status r = scratch_for(st, tiles, entries, &sc);
        if (r == BZ_OK) *out = sc->cs;
        return r;
    };
    Scene* scn = scene(a->scene);
    const bool cache = (a->flags & BZ_CULL_FRAME_CACHE) != 0;
    float key[19];
    std::memcpy(key, job.view.m, sizeof job.view.m);
    key[18] = job.view.lodTarget;
    key[12] = job.view.fR;
    key[13] = job.view.fL;
    key[14] = job.view.fT;
    key[15] = job.view.fB;
    key[16] = job.view.znear;
    key[17] = job.view.zfar;
    Scene::FrameCache& fc = scn->cache;
    CompactScratch cs{};
    if (cache && job.pass == PASS_LATE && fc.valid && fc.begin == job.begin && fc.n == job.n &&
        fc.xf_version == scn->xf_version && std::memcmp(fc.view_key, key, sizeof key) == 0) {
        // the EARLY pass of this view recorded the frustum survivors: read those (and the current visibility)
        // instead of every transform
        job.n_tiles = tiles_for(job.n, fc.block, fc.items);
        s = compact(late_cached_tiles(job.n), late_cached_entries(job.n), &cs);
        if (s) return s;
        const FrameCacheDev d{ fc.sph, fc.meta, fc.counts };
        hipError_t e = launch_late_cached(job, cs, d, fc.block, fc.items, st);
        if (e != hipSuccess) return hip_fail(e, "cached late-pass launch");
        ++fc.reused;
        return BZ_OK;
    }
    s = compact(job.n_tiles, (uint64_t)job.n_tiles * t.block * t.items, &cs);
    if (s) return s;
    if (cache && job.pass == PASS_EARLY) {
        const uint64_t entries = (uint64_t)job.n_tiles * t.block * t.items;
        if (fc.cap_entries < entries || fc.cap_tiles < job.n_tiles) {
            if (fixed)
                return fail(BZ_E_STATE, "the scene's frame cache is not allocated for this range: run one bz_cull "
                                        "with BZ_CULL_FRAME_CACHE first");
            hipError_t e = hipStreamSynchronize(st);
            for (void* p : { (void*)fc.sph, (void*)fc.meta, (void*)fc.counts })
                if (p) (void)hipFree(p);
            fc.sph = nullptr;
            fc.meta = nullptr;
            fc.counts = nullptr;
            fc.cap_entries = 0;
            fc.cap_tiles = 0;
            if (e == hipSuccess) e = hipMalloc((void**)&fc.sph, (size_t)entries * sizeof(float4));
            if (e == hipSuccess) e = hipMalloc((void**)&fc.meta, (size_t)entries * sizeof(uint2));
            if (e == hipSuccess) e = hipMalloc((void**)&fc.counts, (size_t)job.n_tiles * 4);
            if (e != hipSuccess) {
                fc.valid = false;
                return hip_fail(e, "frame cache allocation");
            }
            fc.cap_entries = entries;
            fc.cap_tiles = job.n_tiles;
        }
        fc.valid = true;
        std::memcpy(fc.view_key, key, sizeof key);
        fc.begin = job.begin;
        fc.n = job.n;
        fc.block = t.block;
        fc.items = t.items;
        fc.xf_version = scn->xf_version;
        const FrameCacheDev d{ fc.sph, fc.meta, fc.counts };
        hipError_t e = launch_cull(job, cs, t.block, t.items, n_cu, t.grid_per_cu, st, &d);
        if (e != hipSuccess) {
            fc.valid = false;
            return hip_fail(e, "cull launch");
        }
        ++fc.recorded;
        return BZ_OK;
    }
    hipError_t e = launch_cull(job, cs, t.block, t.items, n_cu, t.grid_per_cu, st);
    if (e != hipSuccess) return hip_fail(e, "cull launch");
    return BZ_OK;
}
}  // namespace bz

bz_status bz_cull(bz_ctx* ctx, const bz_cull_args* a, void* stream)
{
    if (!ctx || !a || !a->view) return BZ_E_INVALID;
    return ctx->cull(a, (hipStream_t)stream, nullptr);
}

// ---------------------------------------------------------------------------- LOD-instanced cull
bz_status bz_cull_instanced(bz_ctx* ctx, const bz_instanced_cull_args* a, void* stream)
{
    if (!ctx || !a || !a->view) return BZ_E_INVALID;
    Scene* s = ctx->scene(a->scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (!a->instances_per_lod) return ctx->fail(BZ_E_INVALID, "instances_per_lod must be > 0");
    if ((uint64_t)s->n_lods * a->instances_per_lod > 0xffffffffull)
        return ctx->fail(BZ_E_INVALID, "n_lods * instances_per_lod exceeds 2^32 - 1");
    if (!a->instance_ids_dev || !a->counters_dev || !a->count_dev || (a->capacity && !a->cmds_dev))
        return ctx->fail(BZ_E_INVALID, "NULL output buffer");
    if ((uintptr_t)a->cmds_dev % 16u) return ctx->fail(BZ_E_INVALID, "cmds_dev must be 16-byte aligned");
    const bz_tuning& t = ctx->tune[0];
    CullJob job;
    bz_status st_ = ctx->make_job(PASS_PLAIN, 0, a->scene, 0, *a->view, 0, 0, nullptr, nullptr, nullptr, a->count_dev,
                                  0, t.block, t.items, &job);
    if (st_) return st_;
    hipStream_t st = (hipStream_t)stream;
    InstancedJob ij;
    std::memset(&ij, 0, sizeof ij);
    ij.lod_draw = s->lod_draw;
    ij.instance_ids = a->instance_ids_dev;
    ij.counters = reinterpret_cast<uint32_t*>(a->counters_dev);
    ij.cmds = reinterpret_cast<uint32_t*>(a->cmds_dev);
    ij.count = a->count_dev;
    ij.n_lods = s->n_lods;
    ij.per_lod = a->instances_per_lod;
    ij.capacity = a->capacity;
    ij.tile_size = t.block * t.items;
    if (s->n_lods == 0) {
        hipError_t e = launch_fill_u32(a->count_dev, 0u, 1, st);
        return e == hipSuccess ? BZ_OK : ctx->hip_fail(e, "count fill");
    }
    StreamScratch* sc = nullptr;
    st_ = ctx->scratch_for(st, job.n_tiles, (uint64_t)job.n_tiles * t.block * t.items, &sc);
    if (st_) return st_;
    // Chunks of decided tiles, each with a histogram row of n_lods counters. When the LODs fit the decide kernel's
    // LDS histogram, a chunk is one decide workgroup's contiguous run of tiles (as many chunks as workgroups fit at
    // once); otherwise the histogram kernel counts up to 512 chunks (at most ~16 M counters).
    const bool fused = s->n_lods <= kInstLdsBins;
    if (job.n) {
        ij.n_tiles = job.n_tiles;
        if (fused) {
            decide_inst_chunks(job, t.block, t.items, ctx->n_cu, t.grid_per_cu, &ij.tiles_per_chunk, &ij.n_long,
                               &ij.n_chunks);
        } else {
            const uint32_t target = std::max<uint32_t>(1u, std::min<uint32_t>(512u, (1u << 24) / s->n_lods));
            ij.tiles_per_chunk = (job.n_tiles + target - 1) / target;
            ij.n_long = 0;
            ij.n_chunks = (job.n_tiles + ij.tiles_per_chunk - 1) / ij.tiles_per_chunk;
        }
    }
    const uint64_t need = std::max<uint64_t>(1, (uint64_t)ij.n_chunks * s->n_lods);
    const uint64_t need_seg = inst_seg_words(ij.n_chunks, s->n_lods);
    if (sc->inst_hist_cap < need || sc->inst_seg_cap < need_seg) {
        hipError_t e = hipStreamSynchronize(st);
        if (sc->inst_hist_cap < need) {
            if (sc->inst_hist) (void)hipFree(sc->inst_hist);
            sc->inst_hist = nullptr;
            sc->inst_hist_cap = 0;
            if (e == hipSuccess) e = hipMalloc((void**)&sc->inst_hist, need * 4);
            if (e == hipSuccess) sc->inst_hist_cap = need;
        }
        if (e == hipSuccess && sc->inst_seg_cap < need_seg) {
            // the scan's published totals: zero = not yet published
            if (sc->inst_seg) (void)hipFree(sc->inst_seg);
            sc->inst_seg = nullptr;
            sc->inst_seg_cap = 0;
            e = hipMalloc((void**)&sc->inst_seg, need_seg * 8);
            if (e == hipSuccess) e = hipMemset(sc->inst_seg, 0, need_seg * 8);
            if (e == hipSuccess) sc->inst_seg_cap = need_seg;
        }
        if (e != hipSuccess) return ctx->hip_fail(e, "instancing scratch");
    }
    if (sc->inst_seg_dirty) {
        hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(sc->inst_seg), 0u, sc->inst_seg_cap * 2, st);
        if (e != hipSuccess) return ctx->hip_fail(e, "instancing scratch clear");
        sc->inst_seg_dirty = false;
    }
    ij.err = sc->err;
    ij.spin_limit = ctx->wait_limit;
    if (job.n) {
        const InstHistDev ih{ sc->inst_hist, ij.tiles_per_chunk, ij.n_long, ij.n_lods };
        hipError_t e = launch_decide_only(job, sc->cs, t.block, t.items, ctx->n_cu, t.grid_per_cu, st,
                                          fused ? &ih : nullptr);
        if (e != hipSuccess) return ctx->hip_fail(e, "decide launch");
    }
    sc->inst_seg_dirty = true;
    hipError_t e = launch_instanced(ij, sc->cs, sc->inst_hist, sc->inst_seg, fused, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "instanced launch");
    sc->inst_seg_dirty = false;
    return BZ_OK;
}

// ---------------------------------------------------------------------------- cluster cull
bz_status bz_cull_clusters(bz_ctx* ctx, const bz_cluster_cull_args* a, void* stream)
{
    if (!ctx || !a || !a->view) return BZ_E_INVALID;
    Scene* s = ctx->scene(a->scene);
    if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
    if (!s->cl_sphere) return ctx->fail(BZ_E_STATE, "the scene has no cluster table (bz_scene_set_clusters)");
    if (a->cluster_tests & ~(BZ_CLUSTER_FRUSTUM | BZ_CLUSTER_HIZ | BZ_CLUSTER_CONE))
        return ctx->fail(BZ_E_INVALID, "unknown cluster test bits");
    if (a->flags & ~BZ_CLUSTER_OBJECT_HIZ) return ctx->fail(BZ_E_INVALID, "unknown cluster cull flags");
    if ((a->flags & BZ_CLUSTER_OBJECT_HIZ) && !a->pyramid)
        return ctx->fail(BZ_E_INVALID, "BZ_CLUSTER_OBJECT_HIZ needs a pyramid");
    if (!a->dispatch_count_dev || !a->count_dev) return ctx->fail(BZ_E_INVALID, "NULL count buffer");
    if (a->dispatch_capacity && !a->dispatch_dev) return ctx->fail(BZ_E_INVALID, "NULL dispatch buffer");
    if ((uintptr_t)a->dispatch_dev % 4u) return ctx->fail(BZ_E_INVALID, "dispatch_dev must be 4-byte aligned");
    if ((uint64_t)s->n_objects * s->max_lod_clusters > 0xffffffffull)
        return ctx->fail(BZ_E_INVALID, "the scene can expand to more than 2^32-1 cluster records");
    const bz_tuning& t = ctx->tune[0];
    if (t.block * t.items > 2048u) return ctx->fail(BZ_E_INVALID, "object tile larger than 2048");
    // the object pass: PreClusterDrawCull's frustum + LOD (PLAIN), or with the object Hi-Z test (TEMPORAL)
    const bool obj_hiz = (a->flags & BZ_CLUSTER_OBJECT_HIZ) != 0;
    CullJob oj;
    bz_status st_ = ctx->make_job(obj_hiz ? PASS_TEMPORAL : PASS_PLAIN, 0, a->scene, obj_hiz ? a->pyramid : 0,
                                  *a->view, 0, 0, nullptr, nullptr, a->cmds_dev, a->count_dev, a->capacity, t.block,
                                  t.items, &oj);
    if (st_) return st_;
    PyramidDesc cl_pyr{};  // the cluster pass's Hi-Z pyramid
    if (a->pyramid) {
        Pyramid* p = ctx->pyramid(a->pyramid);
        if (!p) return ctx->fail(BZ_E_INVALID, "unknown pyramid handle");
        cl_pyr = p->desc();
    }
    hipStream_t st = (hipStream_t)stream;
    if (oj.n == 0) {
        hipError_t e = launch_fill_u32(a->dispatch_count_dev, 0u, 1, st);
        if (e == hipSuccess) e = launch_fill_u32(a->count_dev, 0u, 1, st);
        if (e != hipSuccess) return ctx->hip_fail(e, "count fill");
        return BZ_OK;
    }
    StreamScratch* sc = nullptr;
    st_ = ctx->scratch_for(st, oj.n_tiles, (uint64_t)oj.n_tiles * t.block * t.items, &sc);
    if (st_) return st_;
    st_ = ctx->cluster_scratch_for(st, sc, oj.n_tiles, a->dispatch_capacity);
    if (st_) return st_;
    ClusterJob cj;
    std::memset(&cj, 0, sizeof cj);
    cj.scene = oj.scene;
    cj.view = oj.view;
    cj.pyr = cl_pyr;
    cj.cl_sphere = s->cl_sphere;
    cj.cl_cone = s->cl_cone;
    cj.cl_draw = s->cl_draw;
    cj.dispatch = reinterpret_cast<uint32_t*>(a->dispatch_dev);
    cj.dispatch_count = a->dispatch_count_dev;
    cj.cmds = reinterpret_cast<uint32_t*>(a->cmds_dev);
    cj.count = a->count_dev;
    cj.dispatch_capacity = a->dispatch_capacity;
    cj.capacity = a->capacity;
    cj.tests = a->cluster_tests;
    cj.hiz = a->pyramid ? 1u : 0u;
    cj.n_tiles = oj.n_tiles;
    cj.tile_size = t.block * t.items;
    hipError_t e = launch_cluster_objects(oj, sc->cs, sc->xs.ccounts, t.block, t.items, st);
    if (e == hipSuccess) e = launch_cluster_pass(cj, sc->cs, sc->xs, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "cluster cull launch");
    return BZ_OK;
}

namespace bz {
bz_status Ctx::upload_jobs(StreamScratch* sc, const CullJob* jobs, uint32_t n, hipStream_t st)
{
    std::vector<uint32_t> starts(n);
    for (uint32_t k = 0; k < n; ++k) starts[k] = jobs[k].tile_start;
    return upload_table(this, sc->jobs_dev, sc->jobs_host, sc->jobs_cap, sc->upload_done, jobs, starts.data(), n, st,
                        "job table upload");
}

bz_status Ctx::upload_pjobs(StreamScratch* sc, const PyrBuildJob* jobs, uint32_t n, hipStream_t st)
{
    std::vector<uint32_t> starts(n);
    for (uint32_t k = 0; k < n; ++k) starts[k] = jobs[k].wg_start;
    return upload_table(this, sc->pjobs_dev, sc->pjobs_host, sc->pjobs_cap, sc->upload_done, jobs, starts.data(), n,
                        st, "pyramid job table upload");
}

bz_status Ctx::make_batch(const bz_view_job* jobs, uint32_t n_jobs, const bz_tuning& t, std::vector<CullJob>& out,
                          std::vector<uint32_t*>& empty, uint32_t* total_tiles)
{
    out.clear();
    out.reserve(n_jobs);
    empty.clear();
    uint64_t tiles = 0;
    for (uint32_t k = 0; k < n_jobs; ++k) {
        const bz_view_job& v = jobs[k];
        if (v.pass != PASS_LATE && v.pass != PASS_PLAIN && v.pass != PASS_TEMPORAL)
            return fail(BZ_E_INVALID, "batched jobs support the LATE, PLAIN and TEMPORAL passes");
        CullJob j;
        bz_status s = make_job(v.pass, 0, v.scene, v.pyramid, v.view, v.object_begin, v.object_count,
                               v.visibility_in_dev, v.visibility_out_dev, v.cmds_dev, v.count_dev, v.capacity, t.block,
                               t.items, &j);
        if (s) return s;
        if (j.n == 0) {  // the kernels assume every job has objects; an empty one only gets its count zeroed
            empty.push_back(j.count);
            continue;
        }
        j.tile_start = (uint32_t)tiles;
        tiles += j.n_tiles;
        if (tiles > 0x7fffffffull) return fail(BZ_E_INVALID, "batch too large");
        out.push_back(j);
    }
    *total_tiles = (uint32_t)tiles;
    return BZ_OK;
}

static hipError_t zero_counts(const std::vector<uint32_t*>& empty, hipStream_t st)
{
    hipError_t e = hipSuccess;
    for (uint32_t* c : empty)
        if (e == hipSuccess) e = launch_fill_u32(c, 0u, 1, st);
    return e;
}

uint32_t Ctx::batch_grid(const bz_tuning& t, uint32_t total_tiles) const
{
    uint64_t g = (uint64_t)n_cu * std::max<uint32_t>(t.grid_per_cu, 1u);
    if (g > total_tiles) g = total_tiles;
    return g ? (uint32_t)g : 1u;
}
}  // namespace bz

bz_status bz_cull_batch(bz_ctx* ctx, const bz_view_job* jobs, uint32_t n_jobs, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    if (n_jobs == 0) return BZ_OK;
    if (!jobs) return ctx->fail(BZ_E_INVALID, "NULL jobs");
    const bz_tuning t = ctx->tune[1];
    std::vector<CullJob> cj;
    std::vector<uint32_t*> empty;
    uint32_t tiles = 0;
    bz_status s = ctx->make_batch(jobs, n_jobs, t, cj, empty, &tiles);
    if (s) return s;
    hipStream_t st = (hipStream_t)stream;
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return ctx->fail(BZ_E_STATE, "bz_cull_batch cannot be stream-captured; use bz_batch_graph_create");
    }
    hipError_t e = zero_counts(empty, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "count fill");
    if (cj.empty()) return BZ_OK;
    StreamScratch* sc = nullptr;
    s = ctx->scratch_for(st, tiles, (uint64_t)tiles * t.block * t.items, &sc);
    if (s) return s;
    s = ctx->upload_jobs(sc, cj.data(), (uint32_t)cj.size(), st);
    if (s) return s;
    e = launch_cull_batch(cj.data(), sc->jobs_dev, table_starts(sc->jobs_dev, sc->jobs_cap), (uint32_t)cj.size(),
                          tiles, max_job_tiles(cj.data(), cj.size()), sc->cs, t.block, t.items, st);
    if (e == hipSuccess) e = hipEventRecord(sc->upload_done, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "batched cull launch");
    return BZ_OK;
}

bz_status bz_get_tuning(bz_ctx* ctx, uint32_t kernel, bz_tuning* out)
{
    if (!ctx || !out || kernel > 1) return BZ_E_INVALID;
    *out = ctx->tune[kernel];
    return BZ_OK;
}

bz_status bz_set_tuning(bz_ctx* ctx, uint32_t kernel, const bz_tuning* t)
{
    if (!ctx || !t || kernel > 1) return BZ_E_INVALID;
    const uint32_t b = t->block, i = t->items;
    const bool ok = (b == 256 || b == 512) && (i == 2 || i == 4);
    if (!ok) return ctx->fail(BZ_E_INVALID, "unsupported tuning: block 256 or 512, items 2 or 4");
    ctx->tune[kernel] = *t;
    return BZ_OK;
}

bz_status bz_check(bz_ctx* ctx, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    // the flags of the stream's own launches and of the graphs last launched on it, each cleared once read
    uint32_t flags[4] = { 0, 0, 0, 0 };
    auto read_flags = [&](uint32_t* err) -> hipError_t {
        uint32_t f[4] = { 0, 0, 0, 0 };
        hipError_t r = hipMemcpy(f, err, sizeof f, hipMemcpyDeviceToHost);
        if (r != hipSuccess) return r;
        if (f[0] | f[1] | f[2] | f[3]) r = hipMemset(err, 0, sizeof f);
        for (int i = 0; i < 4; ++i) flags[i] |= f[i];
        return r;
    };
    auto it = ctx->scratch.find((uintptr_t)st);
    if (it != ctx->scratch.end() && it->second->err) e = read_flags(it->second->err);
    for (auto& g : ctx->graphs)
        if (e == hipSuccess && g.second->launched && g.second->last_stream == st) e = read_flags(g.second->err);
    if (e != hipSuccess) return ctx->hip_fail(e, "error flag read");
    if (flags[1]) {
        // A tail workgroup that gave up may have missed hand-off words that tiles stored after it looked: clear every
        // pyramid's words (with the device idle, so no build is in flight) for the next builds to start from zero.
        e = hipDeviceSynchronize();
        for (auto& p : ctx->pyramids)
            if (e == hipSuccess && p.second->tail)
                e = hipMemset(p.second->tail, 0, (size_t)p.second->tiles_x * p.second->tiles_y * 8u);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) return ctx->hip_fail(e, "pyramid hand-off reset");
    }
    if (flags[0] || flags[1] || flags[2]) {
        if (flags[2] && !flags[0] && !flags[1])
            return ctx->fail(BZ_E_INVALID, "a staged transform update named a transform id out of range (skipped)");
        return ctx->fail(BZ_E_INTERNAL, "a kernel gave up a bounded wait (the pyramid built then holds wrong values "
                                        "until rebuilt)");
    }
    return BZ_OK;
}

bz_status bz_set_wait_limit(bz_ctx* ctx, uint32_t sleeps)
{
    if (!ctx) return BZ_E_INVALID;
    ctx->wait_limit = sleeps;
    return BZ_OK;
}

// ---------------------------------------------------------------------------- graphs
bz_status bz_batch_graph_create(bz_ctx* ctx, const bz_pyramid_job* pyr_jobs, uint32_t n_pyr, const bz_view_job* jobs,
                                uint32_t n_jobs, bz_graph* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    *out = 0;
    if (n_jobs && !jobs) return ctx->fail(BZ_E_INVALID, "NULL jobs");
    if (n_pyr && !pyr_jobs) return ctx->fail(BZ_E_INVALID, "NULL pyramid jobs");
    std::unique_ptr<Graph> g(new (std::nothrow) Graph());
    if (!g) return BZ_E_OOM;
    const bz_tuning t = ctx->tune[1];
    std::vector<CullJob> cj;
    std::vector<uint32_t*> empty;
    uint32_t tiles = 0;
    bz_status s = n_jobs ? ctx->make_batch(jobs, n_jobs, t, cj, empty, &tiles) : BZ_OK;
    if (s) return s;
    std::vector<PyrBuildJob> pj;
    std::vector<std::pair<Pyramid*, bz_pyramid_job>> multi;
    uint32_t wg = 0;
    for (uint32_t k = 0; k < n_pyr; ++k) {
        Pyramid* p = ctx->pyramid(pyr_jobs[k].pyramid);
        if (!p) return ctx->fail(BZ_E_INVALID, "unknown pyramid handle");
        if (!p->single_pass) {
            multi.push_back({ p, pyr_jobs[k] });
            continue;
        }
        PyrBuildJob j = p->build_job(pyr_jobs[k].depth_dev, pyr_jobs[k].w, pyr_jobs[k].h);
        j.spin_limit = ctx->wait_limit;
        j.wg_start = wg;
        wg += j.tiles_x * j.tiles_y;
        pj.push_back(j);
    }
    // graph-owned device state
    hipError_t e = hipSuccess;
    const uint32_t cap = std::max<uint32_t>(tiles, 1u);
    e = hipMalloc((void**)&g->cs.counts, (size_t)counts_words(cap) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&g->cs.staging, (size_t)cap * t.block * t.items * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc((void**)&g->err, 16);
    if (e == hipSuccess) e = hipMemset(g->err, 0, 16);
    if (e != hipSuccess) return ctx->hip_fail(e, "graph scratch");
    g->cs.capacity_tiles = cap;
    g->cs.capacity_entries = (uint64_t)cap * t.block * t.items;
    set_prefix(g->cs);
    // job tables (descriptors, then the compact start indices), owned by the graph
    if (!cj.empty()) {
        std::vector<uint32_t> starts(cj.size());
        for (size_t k = 0; k < cj.size(); ++k) starts[k] = cj[k].tile_start;
        e = hipMalloc((void**)&g->jobs_dev, cj.size() * (sizeof(CullJob) + 4));
        if (e == hipSuccess) e = hipMemcpy(g->jobs_dev, cj.data(), cj.size() * sizeof(CullJob), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(const_cast<uint32_t*>(table_starts(g->jobs_dev, (uint32_t)cj.size())), starts.data(),
                          starts.size() * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !pj.empty()) {
        std::vector<uint32_t> starts(pj.size());
        for (size_t k = 0; k < pj.size(); ++k) starts[k] = pj[k].wg_start;
        e = hipMalloc((void**)&g->pjobs_dev, pj.size() * (sizeof(PyrBuildJob) + 4));
        if (e == hipSuccess)
            e = hipMemcpy(g->pjobs_dev, pj.data(), pj.size() * sizeof(PyrBuildJob), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(const_cast<uint32_t*>(table_starts(g->pjobs_dev, (uint32_t)pj.size())), starts.data(),
                          starts.size() * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "graph job tables");
    hipStream_t cs = nullptr;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return ctx->hip_fail(e, "capture stream");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess && !pj.empty())
    {
        uint32_t mt = 0;
        for (auto& q : pj) mt = std::max<uint32_t>(mt, q.tiles_x * q.tiles_y);
        e = launch_pyramid_batch(g->pjobs_dev, (uint32_t)pj.size(), mt, g->err + 1, cs);
    }
    for (auto& m : multi) {
        if (e != hipSuccess) break;
        Pyramid* p = m.first;
        e = launch_pyramid_level(m.second.depth_dev, m.second.w, m.second.h, p->texels, p->base_w, p->base_h, cs);
        for (uint32_t L = 1; L < p->mips && e == hipSuccess; ++L)
            e = launch_pyramid_level(p->texels + p->level_offset[L - 1], level_dim(p->base_w, L - 1),
                                     level_dim(p->base_h, L - 1), p->texels + p->level_offset[L],
                                     level_dim(p->base_w, L), level_dim(p->base_h, L), cs);
    }
    if (e == hipSuccess) e = zero_counts(empty, cs);
    if (e == hipSuccess && !cj.empty())
        e = launch_cull_batch(cj.data(), g->jobs_dev, table_starts(g->jobs_dev, (uint32_t)cj.size()),
                              (uint32_t)cj.size(), tiles, max_job_tiles(cj.data(), cj.size()), g->cs, t.block,
                              t.items, cs);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(cs, &graph);
    (void)hipStreamDestroy(cs);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return ctx->hip_fail(e, "graph capture");
    const bz_graph h = ++ctx->next_handle;
    ctx->graphs[h] = std::move(g);
    *out = h;
    return BZ_OK;
}

bz_status bz_graph_launch(bz_ctx* ctx, bz_graph graph, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->graphs.find(graph);
    if (it == ctx->graphs.end()) return ctx->fail(BZ_E_INVALID, "unknown graph handle");
    hipError_t e = hipGraphLaunch(it->second->exec, (hipStream_t)stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipGraphLaunch");
    it->second->last_stream = (hipStream_t)stream;
    it->second->launched = true;
    return BZ_OK;
}

bz_status bz_graph_destroy(bz_ctx* ctx, bz_graph graph)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->graphs.find(graph);
    if (it == ctx->graphs.end()) return ctx->fail(BZ_E_INVALID, "unknown graph handle");
    (void)hipDeviceSynchronize();
    ctx->graphs.erase(it);
    return BZ_OK;
}

namespace bz {
Graph::~Graph()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    void* ps[] = { cs.counts, cs.staging, err, jobs_dev, pjobs_dev };
    for (void* p : ps)
        if (p) (void)hipFree(p);
}
}  // namespace bz

// ---------------------------------------------------------------------------- device plumbing
bz_status bz_dev_alloc(bz_ctx* ctx, uint64_t bytes, void** out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    *out = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipMalloc");
    return BZ_OK;
}

bz_status bz_dev_free(bz_ctx* ctx, void* p)
{
    if (!ctx) return BZ_E_INVALID;
    if (!p) return BZ_OK;
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipFree");
    return BZ_OK;
}

static bz_status copy(bz_ctx* ctx, void* dst, const void* src, uint64_t bytes, hipMemcpyKind k, void* stream,
                      bool sync)
{
    if (!ctx || (bytes && (!dst || !src))) return BZ_E_INVALID;
    if (!bytes) return BZ_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream);
    if (e == hipSuccess && sync) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipMemcpyAsync");
    return BZ_OK;
}

bz_status bz_memcpy_h2d(bz_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream)
{
    return copy(ctx, dst, src, bytes, hipMemcpyHostToDevice, stream, true);
}
bz_status bz_memcpy_d2h(bz_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream)
{
    return copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost, stream, true);
}
bz_status bz_memcpy_d2d(bz_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream)
{
    return copy(ctx, dst, src, bytes, hipMemcpyDeviceToDevice, stream, false);
}

bz_status bz_memset_d32(bz_ctx* ctx, void* dst, uint32_t value, uint64_t count, void* stream)
{
    if (!ctx || (count && !dst)) return BZ_E_INVALID;
    hipError_t e = launch_fill_u32((uint32_t*)dst, value, count, (hipStream_t)stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "fill");
    return BZ_OK;
}

bz_status bz_stream_create(bz_ctx* ctx, void** stream)
{
    if (!ctx || !stream) return BZ_E_INVALID;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamCreate");
    *stream = s;
    return BZ_OK;
}

bz_status bz_stream_destroy(bz_ctx* ctx, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    if (!stream) return BZ_OK;
    (void)hipStreamSynchronize((hipStream_t)stream);
    ctx->scratch.erase((uintptr_t)stream);
    hipError_t e = hipStreamDestroy((hipStream_t)stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamDestroy");
    return BZ_OK;
}

bz_status bz_stream_sync(bz_ctx* ctx, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipStreamSynchronize");
    return BZ_OK;
}

bz_status bz_synth_depth_device(bz_ctx* ctx, uint64_t seed, uint32_t w, uint32_t h, uint32_t occluders, float p11,
                                float z_near, float* depth_dev, void* stream)
{
    if (!ctx || !depth_dev || !w || !h) return BZ_E_INVALID;
    if (occluders > 4096) return ctx->fail(BZ_E_INVALID, "at most 4096 occluders");
    hipError_t e = launch_synth_depth(seed, w, h, occluders, p11, z_near, depth_dev, (hipStream_t)stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "synthetic depth launch");
    return BZ_OK;
}

}  // extern "C"
