// cull.hip — gfx950 draw-cull kernels (early / late / plain / temporal / onpc passes).
//
// Replaces InitialDrawCull.comp.glsl:12-58, LateDrawCull.comp.glsl:13-71, TransparentDrawCull.comp.glsl:13-65,
// OnpcDrawCull.comp.glsl:11-52 and drawOccTemporal.hlsl:13-64. The per-object arithmetic is in cull_math.h.
//
// Two kernels per pass, both HBM-streaming, with no wait between workgroups:
//   decide: one workgroup per tile of BLOCK x ITEMS objects (object tile*TILE + k*BLOCK + tid). Every object stream
//           of the tile is loaded at once (coalesced SoA float4 / uint32 loads); sphere -> view space and frustum
//           test for every object; the frustum survivors of each wave are compacted into LDS so the Hi-Z test
//           (two square roots, five divisions, a 2x2 gather) runs on dense lanes, and likewise the emitting objects
//           for LOD selection. Emitting objects are ranked with a wave64 ballot and an LDS scan in (item, wave)
//           order, i.e. ascending object order, and written as (objectId, LOD) to the tile's staging slot; the
//           tile's count goes to counts[tile]. The LATE pass also writes visibility here.
//   emit:   each workgroup takes a run of tiles, adds the counts of the job's earlier tiles (a block reduction) to
//           get its output base, and writes DrawIndexedIndirect commands in order, plus the draw count.
// The output order is ascending objectId, independent of scheduling (the reference appends in atomicAdd order).
// A single-pass decoupled look-back was measured slower here: cross-XCD status visibility costs ~1 us per round
// trip and the first generation of workgroups walked back tens of round trips.
#include "bz_internal.h"

namespace bz {

namespace {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

// Small tables staged in LDS per workgroup when they fit: surface spheres and LOD ranges (512 surfaces, 12 KiB),
// LOD errors (2048 LODs, 8 KiB) and the pyramid's level offsets. Larger tables are gathered from global memory.
constexpr uint32_t kLdsSurfaces = 512;
constexpr uint32_t kLdsLods = 2048;

template <int BLOCK, int ITEMS>
struct TileSmem {
    float4 surf[kLdsSurfaces];
    uint2 surf_lod[kLdsSurfaces];
    float lod_err[kLdsLods];
    float4 cand[BLOCK];               // per-wave compaction area: one chunk of 64 entries per wave
    uint2 cand2[BLOCK];
    uint32_t lvl[kMaxMips];
    uint32_t woff[ITEMS][BLOCK / 64];
    uint32_t csum[BLOCK / 64];        // cluster mode: per-wave sums of the emitting objects' cluster counts
    uint32_t cwoff[ITEMS][BLOCK / 64];  // frame cache: survivor offsets in (item, wave) order
    uint32_t job;
};

// Pass families compiled separately so the visibility stream is loaded unconditionally where it is used.
enum : int { MODE_EARLY = 0, MODE_LATE = 1, MODE_OTHER = 2 };  // OTHER: PLAIN, TEMPORAL, ONPC (runtime-uniform)

// Orders a wave's LDS writes before its later LDS reads (the compaction areas are private to a wave): a compiler
// barrier, then a wait on LDS operations only (lgkmcnt(0); vmcnt and expcnt left at their maxima), so global loads in
// flight, such as the next tile's prefetch, are not waited on.
__device__ __forceinline__ void wave_lds_sync()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Per-wave compaction of a lane predicate over ITEMS items: slot[k] = this lane's position among the wave's set
// entries in (item, lane) order; returns the wave's total.
template <int ITEMS>
__device__ __forceinline__ uint32_t wave_compact(const bool (&pred)[ITEMS], uint32_t (&slot)[ITEMS], uint32_t lane)
{
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const unsigned long long m = __ballot(pred[k]);
        slot[k] = total + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        total += (uint32_t)__popcll(m);
    }
    return total;
}

// Object streams of one tile held in registers. Indices are clamped into the job so every load is issued without
// waiting on another; out-of-range lanes are discarded afterwards.
template <int ITEMS>
struct Streams {
    uint32_t gi[ITEMS], vis[ITEMS], surf[ITEMS];
    bool valid[ITEMS], active[ITEMS];
    float4 ps[ITEMS], q[ITEMS];
};

// Loads of the per-object streams. NT: marked non-temporal (streaming, evicted first), for the passes that read each
// stream once and gather no Hi-Z texels: an EARLY pass recording a frame cache (its LATE pass reads the records, not
// these streams again: 4% faster per frame), PLAIN and ONPC (4%). Without the cache the LATE pass re-reads the same
// streams and finds part of them still in the last-level cache, which the hint would give up (2.5% slower per frame),
// and passes with a Hi-Z test keep their texels cached better without it, so those keep the default policy.
typedef float nt_float4 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ float4 stream_load(const float4* p)
{
    if constexpr (NT) {
        const nt_float4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_float4*>(p));
        return make_float4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}
template <bool NT, class T>
__device__ __forceinline__ T stream_load(const T* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Tile indices and the visibility stream (EARLY, LATE). Jobs always hold at least one object (empty ranges are
// answered on the host), so clamped indices are always valid.
template <int BLOCK, int ITEMS, int MODE, bool NT = false>
__device__ __forceinline__ void issue_vis(const CullJob& J, uint32_t tile_in_job, Streams<ITEMS>& s)
{
    const uint32_t base = tile_in_job * (uint32_t)(BLOCK * ITEMS);
    const uint32_t nmax = J.n ? J.n - 1u : 0u;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = base + (uint32_t)k * BLOCK + threadIdx.x;
        s.valid[k] = i < J.n;
        s.gi[k] = J.begin + (s.valid[k] ? i : nmax);
        s.vis[k] = 0u;
    }
    if (MODE == MODE_EARLY || MODE == MODE_LATE) {  // J.n >= 1: clamped indices are in range
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) s.vis[k] = stream_load<NT>(J.vis_in + s.gi[k]);
    }
}

// The surface-id and transform streams: for every object, or in EARLY only for last-frame-visible ones (which waits
// for the visibility stream first).
// ALL (EARLY recording a frame cache): every object's streams are loaded, not only the last-frame-visible ones.
template <int ITEMS, int MODE, bool XF_ID, bool ALL = false, bool NT = false>
__device__ __forceinline__ void issue_objects(const CullJob& J, Streams<ITEMS>& s)
{
    const SceneDev& S = J.scene;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        s.active[k] = s.valid[k] && (MODE != MODE_EARLY || s.vis[k] != 0u);
        s.surf[k] = 0u;
        s.ps[k] = s.q[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (MODE != MODE_EARLY || ALL || s.active[k]) {
            const uint32_t x = XF_ID ? s.gi[k] : stream_load<NT>(S.obj_xf + s.gi[k]);
            s.surf[k] = S.obj_surf16 ? (uint32_t)stream_load<NT>(S.obj_surf16 + s.gi[k])
                                     : stream_load<NT>(S.obj_surf + s.gi[k]);
            s.ps[k] = stream_load<NT>(S.xf_ps + x);
            s.q[k] = stream_load<NT>(S.xf_q + x);
        }
    }
}

template <int BLOCK, int ITEMS, int MODE, bool XF_ID, bool ALL = false, bool NT = false>
__device__ __forceinline__ void issue_tile(const CullJob& J, uint32_t tile_in_job, Streams<ITEMS>& s)
{
    issue_vis<BLOCK, ITEMS, MODE, NT>(J, tile_in_job, s);
    issue_objects<ITEMS, MODE, XF_ID, ALL, NT>(J, s);
}

// Small tables into LDS (surface spheres and LOD ranges, LOD errors) plus the pyramid's level offsets. All loads of a
// thread are issued before its first LDS write. The caller places the barrier.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void stage_tables(const CullJob& J, TileSmem<BLOCK, ITEMS>& sm, bool lds_tab)
{
    const uint32_t tid = threadIdx.x;
    const SceneDev& S = J.scene;
    if (lds_tab) {
        constexpr int SPT = (kLdsSurfaces + BLOCK - 1) / BLOCK, LPT = (kLdsLods + BLOCK - 1) / BLOCK;
        float4 t_sph[SPT];
        uint2 t_lod[SPT];
        float t_err[LPT];
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            const uint32_t t = tid + (uint32_t)j * BLOCK;
            const uint32_t tt = t < S.n_surfaces ? t : 0u;
            t_sph[j] = S.surf_sphere[tt];
            t_lod[j] = S.surf_lod[tt];
        }
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
            const uint32_t t = tid + (uint32_t)j * BLOCK;
            t_err[j] = S.lod_error[t < S.n_lods ? t : 0u];
        }
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            const uint32_t t = tid + (uint32_t)j * BLOCK;
            if (t < S.n_surfaces) {
                sm.surf[t] = t_sph[j];
                sm.surf_lod[t] = t_lod[j];
            }
        }
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
            const uint32_t t = tid + (uint32_t)j * BLOCK;
            if (t < S.n_lods) sm.lod_err[t] = t_err[j];
        }
    }
    if (tid < (uint32_t)kMaxMips) {  // level offsets from the extent (no load): sum of the smaller levels' areas
        uint32_t off = 0;
        for (uint32_t L = 0; L < tid; ++L) off += level_dim(J.pyr.base_w, L) * level_dim(J.pyr.base_h, L);
        sm.lvl[tid] = off;
    }
}

struct NoPrefetch {
    __device__ __forceinline__ void operator()() const {}
};

// Decides one tile whose streams are in `s` (the decide kernel's body). Called by all threads of the workgroup.
// `prefetch` is invoked once, after the tile's last global gather (the Hi-Z texels), so loads it issues stay in
// flight through the rest of the tile (vmcnt retires in issue order: nothing later waits on them).
// CL (cluster object pass): ccounts[gtile] also receives the sum of the emitting objects' LOD cluster counts.
// CACHE (EARLY with all streams loaded): every frustum survivor of the tile is recorded in the frame cache `fc`.
// INST (LOD-instanced output): each emitting object also adds 1 to its LOD's bin of the workgroup's LDS histogram.
template <int BLOCK, int ITEMS, int MODE, bool HIZ, bool LDS_TAB, bool CL = false, bool CACHE = false, bool INST = false,
          class Prefetch>
__device__ __forceinline__ void process_tile(const CullJob& J, uint32_t gtile, const Streams<ITEMS>& s,
                                             uint32_t* counts, uint2* staging, uint32_t* ccounts,
                                             TileSmem<BLOCK, ITEMS>& sm, Prefetch&& prefetch,
                                             FrameCacheDev fc = FrameCacheDev{}, uint32_t* lds_hist = nullptr)
{
    constexpr bool lds_tab = LDS_TAB;
    constexpr int NW = BLOCK / 64;
    constexpr uint32_t TILE = BLOCK * ITEMS;
    constexpr bool late = MODE == MODE_LATE;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t pass = J.pass;
    const SceneDev& S = J.scene;

    float4 sph[ITEMS];
    if (lds_tab) {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) sph[k] = sm.surf[s.surf[k]];
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) sph[k] = S.surf_sphere[s.surf[k]];
    }

    // ---- sphere -> view space + frustum for every object
    float cx[ITEMS], cy[ITEMS], cz[ITEMS], rad[ITEMS];
    bool vis_new[ITEMS], emit[ITEMS], frus[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        sphere_to_view(J.view, sph[k].x, sph[k].y, sph[k].z, sph[k].w, s.ps[k].x, s.ps[k].y, s.ps[k].z, s.ps[k].w,
                       s.q[k].x, s.q[k].y, s.q[k].z, s.q[k].w, cx[k], cy[k], cz[k], rad[k]);
        if constexpr (CACHE) {
            frus[k] = s.valid[k] && frustum_visible(J.view, cx[k], cy[k], cz[k], rad[k]);
            vis_new[k] = frus[k] && s.active[k];
        } else {
            vis_new[k] = s.active[k] && frustum_visible(J.view, cx[k], cy[k], cz[k], rad[k]);
        }
    }
    float4* cand = sm.cand + wave * 64u;
    uint2* cand2 = sm.cand2 + wave * 64u;

    // ---- Hi-Z on the wave's frustum survivors, compacted onto dense lanes in chunks of 64. The texel gathers are
    // issued by every lane (idle lanes read a clamped, in-bounds texel) and the prefetch is issued right after the
    // first chunk's gathers: vmcnt retires in issue order, so waiting for the gathers does not wait for the prefetch.
    if constexpr (HIZ) {
        uint32_t slot[ITEMS];
        const uint32_t n = wave_compact<ITEMS>(vis_new, slot, lane);
        bool res[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) res[k] = vis_new[k];
        auto chunk = [&](uint32_t c0, auto&& between) {
#pragma unroll
            for (int k = 0; k < ITEMS; ++k)
                if (vis_new[k] && slot[k] - c0 < 64u) cand[slot[k] - c0] = make_float4(cx[k], cy[k], cz[k], rad[k]);
            wave_lds_sync();
            const bool act = c0 + lane < n;
            const float4 c = cand[lane];
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
            const bool test =
                act && project_sphere(c.x, c.y, c.z, c.w, J.view.znear, J.view.p00, J.view.p11, a0, a1, a2, a3);
            const float width = (a2 - a0) * J.view.pyrW;
            const float height = (a3 - a1) * J.view.pyrH;
            const uint32_t L = hiz_level(glsl_max(width, height), J.pyr.mips);
            const uint32_t W = level_dim(J.pyr.base_w, L), H = level_dim(J.pyr.base_h, L);
            const Footprint f = footprint((a0 + a2) * 0.5f, (a1 + a3) * 0.5f, W, H);
            float d00, d01, d10, d11;
            footprint_texels(J.pyr.texels + sm.lvl[L], W, f, d00, d01, d10, d11);
            __builtin_amdgcn_sched_barrier(0);  // keep the gathers ahead of the prefetch in issue order
            between();
            __builtin_amdgcn_sched_barrier(0);
            if (act) {
                const float d = glsl_min(glsl_min(d00, d01), glsl_min(d10, d11));
                // a sphere crossing the near plane is not tested and stays visible
                cand2[lane].x = (test ? (J.view.znear / (c.z - c.w) > d) : true) ? 1u : 0u;
            }
            wave_lds_sync();
#pragma unroll
            for (int k = 0; k < ITEMS; ++k)
                if (vis_new[k] && slot[k] - c0 < 64u) res[k] = cand2[slot[k] - c0].x != 0u;
            wave_lds_sync();
        };
        chunk(0u, prefetch);  // always runs (all lanes idle when n == 0): the prefetch is issued unconditionally
        for (uint32_t c0 = 64u; c0 < n; c0 += 64u) chunk(c0, NoPrefetch{});
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) vis_new[k] = res[k];
    } else {
        prefetch();
    }

    // ---- LOD selection on the wave's emitting objects (with a frame cache: on every frustum survivor, a superset),
    // compacted onto dense lanes in chunks of 64
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) emit[k] = late ? (vis_new[k] && s.vis[k] == 0u) : vis_new[k];
    bool lsel[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) lsel[k] = CACHE ? frus[k] : emit[k];
    uint32_t eslot[ITEMS], lod_abs[ITEMS];
    const uint32_t ne = wave_compact<ITEMS>(lsel, eslot, lane);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) lod_abs[k] = 0u;
    for (uint32_t c0 = 0; c0 < ne; c0 += 64u) {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (lsel[k] && eslot[k] - c0 < 64u) {
                cand[eslot[k] - c0] = make_float4(cx[k], cy[k], cz[k], rad[k]);
                cand2[eslot[k] - c0] = make_uint2(__float_as_uint(s.ps[k].w), s.surf[k]);
            }
        wave_lds_sync();
        if (c0 + lane < ne) {
            const float4 c = cand[lane];
            const uint2 e = cand2[lane];
            const float scale = __uint_as_float(e.x);
            const uint2 lo = lds_tab ? sm.surf_lod[e.y] : S.surf_lod[e.y];
            const float len = sqrtf((c.x * c.x + c.y * c.y) + c.z * c.z);
            const float distance = glsl_max(len - c.w, 0.0f);
            const float threshold = (distance * J.view.lodTarget) / scale;
            uint32_t idx = 0;
            if (lds_tab) {
                for (uint32_t l = 1; l < lo.y; ++l)
                    if (sm.lod_err[lo.x + l] < threshold) idx = l;
            } else {
                for (uint32_t l = 1; l < lo.y; ++l)
                    if (S.lod_error[lo.x + l] < threshold) idx = l;
            }
            cand2[lane].x = (pass == PASS_ONPC) ? idx : idx + lo.x;
        }
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (lsel[k] && eslot[k] - c0 < 64u) lod_abs[k] = cand2[eslot[k] - c0].x;
        wave_lds_sync();
    }

    if constexpr (INST) {  // the LOD histogram of the workgroup's tiles (LDS)
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (emit[k]) atomicAdd(&lds_hist[lod_abs[k]], 1u);
    }

    if constexpr (CL) {  // the tile's cluster total, for the expansion's record offsets
        uint32_t cc = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (emit[k]) cc += S.lod_cluster[lod_abs[k]].y;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) cc += __shfl_xor(cc, off, 64);
        if (lane == 0) sm.csum[wave] = cc;
    }

    // ---- visibility; ranks in (item, wave) order -> staging
    if (late) {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (s.valid[k]) J.vis_out[s.gi[k]] = vis_new[k] ? 1u : 0u;
    }
    uint32_t rank[ITEMS], crank[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const unsigned long long m = __ballot(emit[k]);
        rank[k] = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) sm.woff[k][wave] = (uint32_t)__popcll(m);
        if constexpr (CACHE) {
            const unsigned long long mc = __ballot(frus[k]);
            crank[k] = (uint32_t)__popcll(mc & ((1ull << lane) - 1ull));
            if (lane == 0) sm.cwoff[k][wave] = (uint32_t)__popcll(mc);
        }
    }
    __syncthreads();
    if constexpr (CACHE) {
        if (wave == (uint32_t)NW - 1u) {  // the cache's survivor offsets, scanned by another wave than the commands'
            constexpr uint32_t E = ITEMS * NW;
            const uint32_t c = lane < E ? (&sm.cwoff[0][0])[lane] : 0u;
            const uint32_t incl = wave_incl_scan(c, lane);
            if (lane < E) (&sm.cwoff[0][0])[lane] = incl - c;
            if (lane == 63) fc.counts[gtile] = incl;
        }
    }
    if (wave == 0) {
        constexpr uint32_t E = ITEMS * NW;  // <= 64
        const uint32_t c = lane < E ? (&sm.woff[0][0])[lane] : 0u;
        const uint32_t incl = wave_incl_scan(c, lane);
        if (lane < E) (&sm.woff[0][0])[lane] = incl - c;
        if (lane == 63) counts[gtile] = incl;
        if constexpr (CL) {
            if (lane == 0) {
                uint32_t tot = 0;
#pragma unroll
                for (int w = 0; w < NW; ++w) tot += sm.csum[w];
                ccounts[gtile] = tot;
            }
        }
    }
    __syncthreads();
    uint2* stage = staging + (size_t)gtile * TILE;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
        if (emit[k]) stage[sm.woff[k][wave] + rank[k]] = make_uint2(s.gi[k], lod_abs[k]);
    if constexpr (CACHE) {
        const size_t cb = (size_t)gtile * TILE;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (frus[k]) {
                const uint32_t pos = sm.cwoff[k][wave] + crank[k];
                fc.sph[cb + pos] = make_float4(cx[k], cy[k], cz[k], rad[k]);
                fc.meta[cb + pos] = make_uint2(s.gi[k], lod_abs[k]);
            }
    }
    __syncthreads();  // woff / cwoff are rewritten by the next tile
}

// One tile, loads and decision (non-persistent form, used by the batched kernel and the cluster object pass).
template <int BLOCK, int ITEMS, int MODE, bool HIZ, bool XF_ID, bool LDS_TAB, bool CL = false>
__device__ __forceinline__ void decide_tile(const CullJob& J, uint32_t tile_in_job, uint32_t gtile, uint32_t* counts,
                                            uint2* staging, TileSmem<BLOCK, ITEMS>& sm, uint32_t* ccounts = nullptr)
{
    constexpr bool lds_tab = LDS_TAB;
    constexpr bool NT = MODE == MODE_OTHER && !HIZ;  // PLAIN / ONPC: streamed once, no Hi-Z
    Streams<ITEMS> s;
    issue_vis<BLOCK, ITEMS, MODE, NT>(J, tile_in_job, s);
    if (MODE != MODE_EARLY) issue_objects<ITEMS, MODE, XF_ID, false, NT>(J, s);
    stage_tables<BLOCK, ITEMS>(J, sm, lds_tab);
    if (MODE == MODE_EARLY) issue_objects<ITEMS, MODE, XF_ID, false, NT>(J, s);
    __syncthreads();  // staged tables visible
    process_tile<BLOCK, ITEMS, MODE, HIZ, LDS_TAB, CL>(J, gtile, s, counts, staging, ccounts, sm, NoPrefetch{});
}

// Persistent decide for one view: workgroup b takes tiles b, b + G, b + 2G, ... (no workgroup waits on another, so
// any residency is safe). Tables are staged once; the next tile's streams are issued while the current one is
// decided (two register sets, alternating, so no register copy waits on a load in flight).
// INST: workgroup b takes the contiguous tiles of chunk b instead (chunk_range), counts
// its emitting objects per LOD in LDS and writes the counts to row b of ih.hist.
template <int BLOCK, int ITEMS, int MODE, bool HIZ, bool XF_ID, bool LDS_TAB, bool CACHE = false, bool INST = false>
__global__ __launch_bounds__(BLOCK) void decide_kernel(CullJob J, uint32_t* counts, uint2* staging, FrameCacheDev fc,
                                                       InstHistDev ih)
{
    __shared__ TileSmem<BLOCK, ITEMS> sm;
    __shared__ uint32_t s_hist[INST ? kInstLdsBins : 1];
    constexpr bool lds_tab = LDS_TAB;
    constexpr bool NT = (MODE == MODE_EARLY && CACHE) || (MODE == MODE_OTHER && !HIZ);
    uint32_t t, t_end, step;
    if constexpr (INST) {
        chunk_range(blockIdx.x, ih.tiles_per_chunk, ih.n_long, J.n_tiles, t, t_end);
        step = 1;
        for (uint32_t i = threadIdx.x; i < ih.n_lods; i += BLOCK) s_hist[i] = 0u;  // (ordered by the barrier below)
    } else {
        t = blockIdx.x;
        t_end = J.n_tiles;
        step = gridDim.x;
    }
    Streams<ITEMS> A, B;
    issue_tile<BLOCK, ITEMS, MODE, XF_ID, CACHE, NT>(J, t, A);
    stage_tables<BLOCK, ITEMS>(J, sm, lds_tab);
    __syncthreads();
    while (true) {
        {
            const uint32_t tn = t + step;
            // past the last tile, tile n_tiles is issued: every lane is out of range and clamped to the job's last
            // object, so each stream costs one cache line (discarded). A branch around the loads instead was 2.7 us
            // slower per pass, re-issuing the current tile 2.2 us slower (with non-temporal loads its lines are
            // fetched again).
            process_tile<BLOCK, ITEMS, MODE, HIZ, LDS_TAB, false, CACHE, INST>(J, t, A, counts, staging, nullptr, sm,
                                                                                [&]() {
                issue_tile<BLOCK, ITEMS, MODE, XF_ID, CACHE, NT>(J, tn < t_end ? tn : J.n_tiles, B);
            }, fc, s_hist);
            t = tn;
            if (t >= t_end) break;
        }
        {
            const uint32_t tn = t + step;
            process_tile<BLOCK, ITEMS, MODE, HIZ, LDS_TAB, false, CACHE, INST>(J, t, B, counts, staging, nullptr, sm,
                                                                                [&]() {
                issue_tile<BLOCK, ITEMS, MODE, XF_ID, CACHE, NT>(J, tn < t_end ? tn : J.n_tiles, A);
            }, fc, s_hist);
            t = tn;
            if (t >= t_end) break;
        }
    }
    if constexpr (INST) {
        __syncthreads();
        uint32_t* row = ih.hist + (size_t)blockIdx.x * ih.n_lods;
        for (uint32_t i = threadIdx.x; i < ih.n_lods; i += BLOCK) row[i] = s_hist[i];
    }
}

// Last entry of the ascending table `starts` that is <= t.
__device__ __forceinline__ uint32_t find_job(const uint32_t* starts, uint32_t n_jobs, uint32_t t)
{
    uint32_t lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr uint32_t kLdsJobs = 1024;  // batched: job start table staged in LDS up to this many jobs

// Batched decide: one launch over the (job, tile) work list of many views; workgroup b takes global tile b. The job
// lookup searches the jobs' first tiles: up to kArgJobs jobs they come in the kernel argument `sa` (scalar loads
// from the kernarg segment, no barrier; the 256-view batch decides in 402.8 us against 414.3 us with the LDS copy),
// up to kLdsJobs they are staged in LDS (one coalesced read), beyond that the compact global table is searched. The
// job itself is read in place with scalar loads (its index is wave-uniform), with no LDS copy and barrier: the
// 256-view graph 0.732 ms against 0.742 ms with the copy. The surface and LOD tables are read from global memory
// (they stay in the L2), not staged in LDS: each workgroup decides a single tile, and staging the whole tables for it
// cost more (the 256-view graph 0.78 ms against 0.76 ms).
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void decide_batch_kernel(const CullJob* __restrict__ jobs,
                                                             const uint32_t* __restrict__ starts, uint32_t n_jobs,
                                                             uint32_t* counts, uint2* staging, JobStartsArg sa)
{
    __shared__ TileSmem<BLOCK, ITEMS> sm;
    __shared__ uint32_t s_start[kLdsJobs];
    const uint32_t t = blockIdx.x;
    uint32_t jb;
    if (n_jobs <= kArgJobs) {
        jb = find_job(sa.v, n_jobs, t);
    } else if (n_jobs <= kLdsJobs) {
        for (uint32_t k = threadIdx.x; k < n_jobs; k += BLOCK) s_start[k] = starts[k];
        __syncthreads();
        jb = find_job(s_start, n_jobs, t);
    } else {
        jb = find_job(starts, n_jobs, t);
    }
    const CullJob& sjob = jobs[__builtin_amdgcn_readfirstlane(jb)];
    const uint32_t tt = t - sjob.tile_start;
    const bool xf_id = sjob.scene.obj_xf == nullptr;
    // identity scenes have no transform-index stream (obj_xf == nullptr): XF_ID must follow it in every variant
#define BZ_DECIDE_TILE(M, HZ)                                                                                     \
    if (xf_id) decide_tile<BLOCK, ITEMS, M, HZ, true, false>(sjob, tt, t, counts, staging, sm);                  \
    else decide_tile<BLOCK, ITEMS, M, HZ, false, false>(sjob, tt, t, counts, staging, sm);
    if (sjob.pass == PASS_LATE) {
        BZ_DECIDE_TILE(MODE_LATE, true)
    } else if (sjob.hiz) {
        BZ_DECIDE_TILE(MODE_OTHER, true)
    } else {
        BZ_DECIDE_TILE(MODE_OTHER, false)
    }
#undef BZ_DECIDE_TILE
}

// Cluster object pass (PreClusterDrawCull.comp.glsl:13-47): the PLAIN (or, with a pyramid, TEMPORAL) decision of
// every object, its LOD, and per tile the number of emitting objects (counts), their (objectId, LOD) list (staging)
// and the sum of their LOD cluster counts (ccounts). One workgroup per tile.
template <int BLOCK, int ITEMS, bool HIZ, bool XF_ID, bool LDS_TAB>
__global__ __launch_bounds__(BLOCK) void decide_cluster_kernel(CullJob J, uint32_t* counts, uint2* staging,
                                                               uint32_t* ccounts)
{
    __shared__ TileSmem<BLOCK, ITEMS> sm;
    decide_tile<BLOCK, ITEMS, MODE_OTHER, HIZ, XF_ID, LDS_TAB, true>(J, blockIdx.x, blockIdx.x, counts, staging, sm,
                                                                     ccounts);
}

template <int BLOCK, int ITEMS>
hipError_t launch_cluster_objects_t(const CullJob& job, const CompactScratch& s, uint32_t* ccounts, hipStream_t st)
{
    const bool lds = job.scene.n_surfaces <= kLdsSurfaces && job.scene.n_lods <= kLdsLods;
    const bool xf_id = job.scene.obj_xf == nullptr;
    const dim3 g(job.n_tiles), b(BLOCK);
#define BZ_CL_LAUNCH(HZ)                                                                                           \
    if (xf_id) {                                                                                                   \
        if (lds) hipLaunchKernelGGL((decide_cluster_kernel<BLOCK, ITEMS, HZ, true, true>), g, b, 0, st, job,      \
                                    s.counts, s.staging, ccounts);                                                 \
        else hipLaunchKernelGGL((decide_cluster_kernel<BLOCK, ITEMS, HZ, true, false>), g, b, 0, st, job,         \
                                s.counts, s.staging, ccounts);                                                     \
    } else {                                                                                                       \
        if (lds) hipLaunchKernelGGL((decide_cluster_kernel<BLOCK, ITEMS, HZ, false, true>), g, b, 0, st, job,     \
                                    s.counts, s.staging, ccounts);                                                 \
        else hipLaunchKernelGGL((decide_cluster_kernel<BLOCK, ITEMS, HZ, false, false>), g, b, 0, st, job,        \
                                s.counts, s.staging, ccounts);                                                     \
    }
    if (job.hiz) {
        BZ_CL_LAUNCH(true)
    } else {
        BZ_CL_LAUNCH(false)
    }
#undef BZ_CL_LAUNCH
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------ emit
constexpr uint32_t kEmitBlock = 256;
constexpr uint32_t kEmitTiles = kEmitBlock / 64;  // one tile per wave
constexpr uint32_t kEmitUnroll = 2;               // entries per lane in flight (a tile emits ~70 on average)

struct EmitSmem {
    uint32_t red[kEmitBlock / 64];
    uint32_t cnt[kEmitTiles + 1];     // exclusive prefix of the workgroup's tile counts
    uint32_t job;
};

// A workgroup barrier for LDS only: waits for this wave's LDS operations (lgkmcnt(0)) but not for its global loads in
// flight (vmcnt and expcnt left at their maxima), then s_barrier.
__device__ __forceinline__ void block_lds_sync()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Writes the commands of tiles [t0, t1) of one job (t1 - t0 <= kEmitTiles; tile_size >= 64 * kEmitUnroll). One round
// of loads gets the counts of the job's earlier tiles (reduced to the output base; at most kScanTiles of them), or
// with `scanned` (the exclusive prefix of all counts) two entries of it, the workgroup's own counts and,
// before any count is known, the first 64 * kEmitUnroll staging entries of each wave's tile (in bounds: every tile
// owns tile_size slots); each wave gathers its entries' LOD draw data while the block scan of the counts runs, then
// writes tile t0 + w's commands in order. Longer tiles continue in rounds of kEmitUnroll entries per lane.
__device__ __forceinline__ void emit_tiles(const CullJob& J, uint32_t t0, uint32_t t1, uint32_t tile_size,
                                           const uint32_t* counts, const uint2* staging, const uint32_t* scanned,
                                           EmitSmem& sm)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t nt = t1 - t0;
    const uint32_t tw = t0 + (wave < nt ? wave : 0u);  // this wave's tile (clamped for waves without one)
    const uint2* stage = staging + (size_t)tw * tile_size;
    const uint2* lod_draw = J.scene.lod_draw;
    uint2 ent[kEmitUnroll], d[kEmitUnroll];
#pragma unroll
    for (uint32_t u = 0; u < kEmitUnroll; ++u) ent[u] = stage[u * 64u + lane];  // speculative: checked against c below
    const uint32_t c = wave < nt ? counts[tw] : 0u;
    // the job's earlier tiles: every load of a thread issued before the first add (16 per pass)
    uint32_t part = 0;
    const uint32_t own = tid < nt ? counts[t0 + tid] : 0u;
    if (scanned) {  // one lane takes the difference of the scanned prefixes; the reduction below adds it
        if (tid == 0) part = scanned[t0] - scanned[J.tile_start];
    } else for (uint32_t u0 = J.tile_start; u0 < t0; u0 += 16u * kEmitBlock) {
        uint32_t v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t u = u0 + (uint32_t)j * kEmitBlock + tid;
            v[j] = u < t0 ? counts[u] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) part += v[j];
    }
#pragma unroll
    for (uint32_t u = 0; u < kEmitUnroll; ++u)
        d[u] = u * 64u + lane < c ? lod_draw[ent[u].y] : make_uint2(0u, 0u);  // in flight through the scan
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off, 64);
    if (lane == 0) sm.red[wave] = part;
    if (wave == 0) {
        uint32_t v = own;
#pragma unroll
        for (int off = 1; off < (int)kEmitTiles; off <<= 1) {
            const uint32_t t = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off) v += t;
        }
        if (lane < kEmitTiles) sm.cnt[lane + 1] = v;
        if (lane == 0) sm.cnt[0] = 0;
    }
    block_lds_sync();
    uint32_t prefix = 0;
#pragma unroll
    for (uint32_t w = 0; w < kEmitBlock / 64u; ++w) prefix += sm.red[w];
    if (tid == 0 && t1 == J.tile_start + J.n_tiles) *J.count = prefix + sm.cnt[nt];
    if (wave >= nt) return;
    const uint32_t base = prefix + sm.cnt[wave];
    for (uint32_t e0 = 0; e0 < c; e0 += 64u * kEmitUnroll) {
        if (e0) {
#pragma unroll
            for (uint32_t u = 0; u < kEmitUnroll; ++u) {
                const uint32_t e = e0 + u * 64u + lane;
                ent[u] = stage[e < c ? e : 0u];  // objectId, absolute LOD (entry 0 exists whenever c > 0)
            }
#pragma unroll
            for (uint32_t u = 0; u < kEmitUnroll; ++u) d[u] = lod_draw[ent[u].y];
        }
#pragma unroll
        for (uint32_t u = 0; u < kEmitUnroll; ++u) {
            const uint32_t e = e0 + u * 64u + lane;
            const uint32_t pos = base + e;
            if (e < c && pos < J.capacity) {
                uint2* cmd = reinterpret_cast<uint2*>(J.cmds + (size_t)pos * 6u);
                cmd[0] = make_uint2(ent[u].x, d[u].x);  // objectId, indexCount
                cmd[1] = make_uint2(1u, d[u].y);         // instanceCount, firstIndex
                cmd[2] = make_uint2(0u, 0u);             // vertexOffset, firstInstance
            }
        }
    }
}

__global__ __launch_bounds__(kEmitBlock) void emit_kernel(CullJob J, uint32_t tile_size, const uint32_t* counts,
                                                          const uint2* staging, const uint32_t* prefix)
{
    __shared__ EmitSmem sm;
    const uint32_t t0 = blockIdx.x * kEmitTiles;
    const uint32_t t1 = t0 + kEmitTiles < J.n_tiles ? t0 + kEmitTiles : J.n_tiles;
    emit_tiles(J, t0, t1, tile_size, counts, staging, prefix, sm);
}

// Batched emit: workgroup b covers global tiles [b*kEmitTiles, ...), split at job boundaries.
__global__ __launch_bounds__(kEmitBlock) void emit_batch_kernel(const CullJob* __restrict__ jobs,
                                                                const uint32_t* __restrict__ starts, uint32_t n_jobs,
                                                                uint32_t total_tiles, uint32_t tile_size,
                                                                const uint32_t* counts, const uint2* staging,
                                                                const uint32_t* prefix)
{
    __shared__ EmitSmem sm;
    uint32_t t = blockIdx.x * kEmitTiles;
    const uint32_t t_end = t + kEmitTiles < total_tiles ? t + kEmitTiles : total_tiles;
    while (t < t_end) {
        __syncthreads();
        if (threadIdx.x == 0) sm.job = find_job(starts, n_jobs, t);
        __syncthreads();
        const CullJob& J = jobs[sm.job];
        const uint32_t job_end = J.tile_start + J.n_tiles;
        const uint32_t t1 = job_end < t_end ? job_end : t_end;
        emit_tiles(J, t, t1, tile_size, counts, staging, prefix, sm);
        t = t1;
    }
}

// The emit launch of one job whose decide launch wrote s.counts / s.staging; a job of more than kScanTiles tiles
// first gets its counts' prefix from the scan.
hipError_t launch_emit(const CullJob& job, uint32_t tile_size, const CompactScratch& s, hipStream_t st)
{
    const uint32_t* prefix = nullptr;
    if (job.n_tiles > kScanTiles) {
        const hipError_t e = launch_scan_counts(s.counts, job.n_tiles, s.prefix, st);
        if (e != hipSuccess) return e;
        prefix = s.prefix;
    }
    hipLaunchKernelGGL(emit_kernel, dim3((job.n_tiles + kEmitTiles - 1) / kEmitTiles), dim3(kEmitBlock), 0, st, job,
                       tile_size, (const uint32_t*)s.counts, (const uint2*)s.staging, prefix);
    return hipGetLastError();
}

// LATE pass from a frame cache: the EARLY pass of the same view, transforms and range recorded each tile's frustum
// survivors (view-space sphere, objectId, LOD index) in ascending objectId order. A workgroup reads its survivors
// only: objects outside the frustum are invisible (visibility 0, no command), survivors take the Hi-Z test, and the
// visible ones read their current visibility (as LateDrawCull.comp.glsl:51 does, so any write to it since the EARLY
// pass is seen) and, when newly visible, are emitted with their recorded LOD -- the same arithmetic on the same values
// as the full LATE pass, so the commands, count and visibility are identical.
constexpr uint32_t kLcBlock = 512;  // threads per workgroup of the cached late pass (kSuperTile objects each)

// A workgroup takes a super-tile of kSuperTile objects (kSuperTile / TILE recorded tiles): the survivors of its
// tiles with their LOD and the Hi-Z test; the current visibility of the survivors that pass it (gathered after the
// test: a survivor occluded now is not emitted whatever it was, and a tenth of them pass, so the pass reads a tenth
// of the words: 17.6 us against 18.4 us when every survivor's word was gathered with its texels); then the
// visibility of the whole super-tile (coalesced; 0 outside the frustum) and the newly visible survivors' commands in
// survivor order. No transform, surface or LOD table is read. (A coalesced load of the whole super-tile's visibility
// instead took 8 more registers per thread, which cost a quarter of the resident workgroups: 25.9 us against 14.6 us
// per C1 pass.)
template <int TILE>
__global__ __launch_bounds__(kLcBlock) void late_cached_kernel(CullJob J, FrameCacheDev fc, uint32_t n_rec_tiles,
                                                               uint32_t* counts, uint2* staging)
{
    constexpr uint32_t BLOCK = kLcBlock, NW = BLOCK / 64;
    constexpr uint32_t S = kSuperTile / TILE;            // recorded tiles per super-tile
    constexpr uint32_t OPT = kSuperTile / BLOCK;         // objects per thread
    constexpr uint32_t U = 2;                            // survivors per thread in the first round
    __shared__ uint32_t s_new[kSuperTile / 32];
    __shared__ uint32_t s_lvl[kMaxMips], s_cnt[S + 1], s_woff[U][NW], s_base;
    const uint32_t st = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t base = st * kSuperTile;  // first object of the super-tile, relative to the job's range
    const uint32_t nt = J.n - base < kSuperTile ? J.n - base : kSuperTile;
    const uint32_t obj0 = J.begin + base;
    // speculative first round, issued before the counts are known: entries [0, WIN) of every recorded tile, survivor
    // slot g = k * BLOCK + tid -> (tile g / WIN, entry g % WIN), i.e. ascending objectId
    constexpr uint32_t WIN = BLOCK * U / S;
    float4 sp[U];
    uint2 mt[U];
#pragma unroll
    for (uint32_t k = 0; k < U; ++k) {
        const uint32_t g = k * BLOCK + tid, rt = st * S + g / WIN;
        const size_t slot = (size_t)(rt < n_rec_tiles ? rt : 0u) * TILE + g % WIN;
        sp[k] = fc.sph[slot];
        mt[k] = fc.meta[slot];
    }
    uint32_t cnt = 0;
    if (tid < S) {
        const uint32_t t = st * S + tid;
        cnt = t < n_rec_tiles ? fc.counts[t] : 0u;
    }
    if (wave == 0) {  // prefix of the recorded tiles' counts (S <= 8 lanes)
        uint32_t x = lane < S ? cnt : 0u;
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane < S) s_cnt[lane + 1] = x;
        if (lane == 0) s_cnt[0] = 0;
    }
    if (tid < (uint32_t)kMaxMips) {
        uint32_t off = 0;
        for (uint32_t L = 0; L < tid; ++L) off += level_dim(J.pyr.base_w, L) * level_dim(J.pyr.base_h, L);
        s_lvl[tid] = off;
    }
    for (uint32_t x = tid; x < kSuperTile / 32; x += BLOCK) s_new[x] = 0u;
    if (tid == 0) s_base = 0;
    __syncthreads();
    const uint32_t total = s_cnt[S];
    uint2* stage = staging + (size_t)st * kSuperTile;
    // one round of up to BLOCK x U survivors: Hi-Z test, visibility bits, commands in (k, wave, lane) order after the
    // earlier rounds
    auto round = [&](const float4 (&sp)[U], const uint2 (&mt)[U], const bool (&have)[U]) {
        bool emit[U];
        uint32_t lod_abs[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            emit[k] = false;
            lod_abs[k] = 0u;
            const uint32_t local = mt[k].x - obj0;
            if (have[k] && local < nt) {  // (an entry outside the super-tile can only come from a mismatched cache)
                float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
                const bool test = project_sphere(sp[k].x, sp[k].y, sp[k].z, sp[k].w, J.view.znear, J.view.p00,
                                                 J.view.p11, a0, a1, a2, a3);
                // a sphere crossing the near plane is not tested and stays visible
                const bool vis =
                    test ? occlusion_visible(J.view, J.pyr, s_lvl, a0, a1, a2, a3, sp[k].z, sp[k].w) : true;
                if (vis) atomicOr(&s_new[local >> 5], 1u << (local & 31u));
                // the current visibility (LateDrawCull:51), read only where it decides: for the visible survivors
                emit[k] = vis && J.vis_in[mt[k].x] == 0u;
                lod_abs[k] = mt[k].y;
            }
        }
        uint32_t rank[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const unsigned long long m = __ballot(emit[k]);
            rank[k] = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) s_woff[k][wave] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (wave == 0) {
            constexpr uint32_t E = U * NW;
            const uint32_t c = lane < E ? (&s_woff[0][0])[lane] : 0u;
            const uint32_t incl = wave_incl_scan(c, lane);
            if (lane < E) (&s_woff[0][0])[lane] = s_base + incl - c;
            if (lane == 63) s_base += incl;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < U; ++k)
            if (emit[k]) stage[s_woff[k][wave] + rank[k]] = make_uint2(mt[k].x, lod_abs[k]);
        __syncthreads();  // s_woff is rewritten by the next round
    };
    bool fast = true;  // every recorded tile's survivors fit its window: the speculative round is the whole pass
#pragma unroll
    for (uint32_t q = 0; q < S; ++q) fast = fast && s_cnt[q + 1] - s_cnt[q] <= WIN;
    if (fast) {
        bool have[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const uint32_t g = k * BLOCK + tid, j = g / WIN;
            have[k] = g % WIN < s_cnt[j + 1] - s_cnt[j];
        }
        round(sp, mt, have);
    } else {
        // survivors in rounds of BLOCK x U; g = k * BLOCK + tid is the survivor's rank in the super-tile
        for (uint32_t g0 = 0; g0 < total; g0 += BLOCK * U) {
            float4 sp2[U];
            uint2 mt2[U];
            bool have[U];
#pragma unroll
            for (uint32_t k = 0; k < U; ++k) {
                const uint32_t g = g0 + k * BLOCK + tid;
                have[k] = g < total;
                uint32_t j = 0;  // recorded tile holding survivor g
#pragma unroll
                for (uint32_t q = 1; q < S; ++q) j += g >= s_cnt[q] ? 1u : 0u;
                const uint32_t e = have[k] ? g - s_cnt[j] : 0u;
                const size_t slot = (size_t)(st * S + (have[k] ? j : 0u)) * TILE + e;  // in the cache region
                sp2[k] = fc.sph[slot];
                mt2[k] = fc.meta[slot];
            }
            round(sp2, mt2, have);
        }
    }
    // the visibility is next read by the following frame's EARLY pass: stored non-temporal (1 us faster here)
#pragma unroll
    for (uint32_t k = 0; k < OPT; ++k) {
        const uint32_t i = k * BLOCK + tid;
        if (i < nt) __builtin_nontemporal_store((s_new[i >> 5] >> (i & 31u)) & 1u, J.vis_out + obj0 + i);
    }
    if (tid == 0) counts[st] = s_base;
}

template <int TILE>
hipError_t launch_late_cached_t(const CullJob& job, const CompactScratch& s, const FrameCacheDev& fc,
                                uint32_t n_rec_tiles, hipStream_t st)
{
    hipLaunchKernelGGL((late_cached_kernel<TILE>), dim3(job.n_tiles), dim3(kLcBlock), 0, st, job, fc, n_rec_tiles,
                       s.counts, s.staging);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_emit(job, kSuperTile, s, st);
}

// Persistent grid: `per_cu` workgroups per CU, or (per_cu == 0) as many as are resident at once per CU.
template <int BLOCK, int ITEMS, int MODE, bool HIZ, bool XF_ID, bool LDS_TAB, bool CACHE = false, bool INST = false>
void launch_decide(const CullJob& job, uint32_t n_cu, uint32_t per_cu, uint32_t* counts, uint2* staging, hipStream_t st,
                   const FrameCacheDev& fc = FrameCacheDev{}, const InstHistDev& ih = InstHistDev{})
{
    static int resident = 0;
    if (resident == 0) {
        int r = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &r, decide_kernel<BLOCK, ITEMS, MODE, HIZ, XF_ID, LDS_TAB, CACHE, INST>, BLOCK, 0) != hipSuccess ||
            r < 1)
            r = 1;
        resident = r;
    }
    uint64_t grid = (uint64_t)n_cu * (per_cu ? per_cu : (uint32_t)resident);
    if (grid > job.n_tiles) grid = job.n_tiles;
    if (grid == 0) grid = 1;
    if (INST) grid = job.n_tiles ? (job.n_tiles - ih.n_long) / ih.tiles_per_chunk : 1;  // one chunk per workgroup
    hipLaunchKernelGGL((decide_kernel<BLOCK, ITEMS, MODE, HIZ, XF_ID, LDS_TAB, CACHE, INST>), dim3((uint32_t)grid),
                       dim3(BLOCK), 0, st, job, counts, staging, fc, ih);
}

// Workgroups per CU of the instanced decide (as launch_decide counts them for the persistent grid).
template <int BLOCK, int ITEMS>
uint32_t inst_resident()
{
    static int resident = 0;
    if (resident == 0) {
        int r = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &r, decide_kernel<BLOCK, ITEMS, MODE_OTHER, false, true, true, false, true>, BLOCK, 0) != hipSuccess ||
            r < 1)
            r = 1;
        resident = r;
    }
    return (uint32_t)resident;
}

// The decide of the LOD-instanced output: PLAIN decisions plus the chunks' LOD histograms.
template <int BLOCK, int ITEMS>
hipError_t launch_decide_inst_t(const CullJob& job, const CompactScratch& s, uint32_t n_cu, uint32_t per_cu,
                                const InstHistDev& ih, hipStream_t st)
{
    const bool lds = job.scene.n_surfaces <= kLdsSurfaces && job.scene.n_lods <= kLdsLods;
    const FrameCacheDev none{};
    if (job.scene.obj_xf == nullptr) {
        if (lds) launch_decide<BLOCK, ITEMS, MODE_OTHER, false, true, true, false, true>(job, n_cu, per_cu, s.counts,
                                                                                      s.staging, st, none, ih);
        else launch_decide<BLOCK, ITEMS, MODE_OTHER, false, true, false, false, true>(job, n_cu, per_cu, s.counts,
                                                                                   s.staging, st, none, ih);
    } else {
        if (lds) launch_decide<BLOCK, ITEMS, MODE_OTHER, false, false, true, false, true>(job, n_cu, per_cu, s.counts,
                                                                                       s.staging, st, none, ih);
        else launch_decide<BLOCK, ITEMS, MODE_OTHER, false, false, false, false, true>(job, n_cu, per_cu, s.counts,
                                                                                    s.staging, st, none, ih);
    }
    return hipGetLastError();
}

template <int BLOCK, int ITEMS, int MODE, bool HIZ, bool CACHE = false>
void launch_mode(const CullJob& job, uint32_t n_cu, uint32_t per_cu, uint32_t* counts, uint2* staging, hipStream_t st,
                 const FrameCacheDev& fc = FrameCacheDev{})
{
    // identity scenes have no transform-index stream (obj_xf == nullptr): XF_ID must follow it in every variant
    const bool lds = job.scene.n_surfaces <= kLdsSurfaces && job.scene.n_lods <= kLdsLods;
    const bool xf_id = job.scene.obj_xf == nullptr;
    if (xf_id) {
        if (lds) launch_decide<BLOCK, ITEMS, MODE, HIZ, true, true, CACHE>(job, n_cu, per_cu, counts, staging, st, fc);
        else launch_decide<BLOCK, ITEMS, MODE, HIZ, true, false, CACHE>(job, n_cu, per_cu, counts, staging, st, fc);
    } else {
        if (lds) launch_decide<BLOCK, ITEMS, MODE, HIZ, false, true, CACHE>(job, n_cu, per_cu, counts, staging, st, fc);
        else launch_decide<BLOCK, ITEMS, MODE, HIZ, false, false, CACHE>(job, n_cu, per_cu, counts, staging, st, fc);
    }
}

template <int BLOCK, int ITEMS>
hipError_t launch_cull_t(const CullJob& job, const CompactScratch& s, uint32_t n_cu, uint32_t per_cu, hipStream_t st,
                         const FrameCacheDev* fc, bool with_emit = true)
{
    const bool h = job.hiz != 0;
    if (job.pass == PASS_EARLY && fc) {
        if (h) launch_mode<BLOCK, ITEMS, MODE_EARLY, true, true>(job, n_cu, per_cu, s.counts, s.staging, st, *fc);
        else launch_mode<BLOCK, ITEMS, MODE_EARLY, false, true>(job, n_cu, per_cu, s.counts, s.staging, st, *fc);
    } else if (job.pass == PASS_EARLY) {
        if (h) launch_mode<BLOCK, ITEMS, MODE_EARLY, true>(job, n_cu, per_cu, s.counts, s.staging, st);
        else launch_mode<BLOCK, ITEMS, MODE_EARLY, false>(job, n_cu, per_cu, s.counts, s.staging, st);
    } else if (job.pass == PASS_LATE) {
        launch_mode<BLOCK, ITEMS, MODE_LATE, true>(job, n_cu, per_cu, s.counts, s.staging, st);
    } else {
        if (h) launch_mode<BLOCK, ITEMS, MODE_OTHER, true>(job, n_cu, per_cu, s.counts, s.staging, st);
        else launch_mode<BLOCK, ITEMS, MODE_OTHER, false>(job, n_cu, per_cu, s.counts, s.staging, st);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !with_emit) return e;
    return launch_emit(job, (uint32_t)(BLOCK * ITEMS), s, st);
}

template <int BLOCK, int ITEMS>
hipError_t launch_cull_batch_t(const CullJob* jobs_host, const CullJob* jobs, const uint32_t* starts, uint32_t n_jobs,
                               uint32_t total_tiles, uint32_t max_job_tiles, const CompactScratch& s, hipStream_t st)
{
    // entries past n_jobs are never searched; they are filled so the argument is fully defined
    JobStartsArg sa;
    for (uint32_t k = 0; k < kArgJobs; ++k) sa.v[k] = k < n_jobs ? jobs_host[k].tile_start : 0xffffffffu;
    hipLaunchKernelGGL((decide_batch_kernel<BLOCK, ITEMS>), dim3(total_tiles), dim3(BLOCK), 0, st, jobs, starts,
                       n_jobs, s.counts, s.staging, sa);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t* prefix = nullptr;
    if (max_job_tiles > kScanTiles) {  // the batch's global prefix; each job takes differences of it
        e = launch_scan_counts(s.counts, total_tiles, s.prefix, st);
        if (e != hipSuccess) return e;
        prefix = s.prefix;
    }
    hipLaunchKernelGGL(emit_batch_kernel, dim3((total_tiles + kEmitTiles - 1) / kEmitTiles), dim3(kEmitBlock), 0, st,
                       jobs, starts, n_jobs, total_tiles, (uint32_t)(BLOCK * ITEMS), (const uint32_t*)s.counts,
                       (const uint2*)s.staging, prefix);
    return hipGetLastError();
}

}  // namespace

#define BZ_TUNING_TABLE(X) \
    X(256, 2) X(256, 4) X(512, 2) X(512, 4)

hipError_t launch_cull(const CullJob& job, const CompactScratch& s, uint32_t block, uint32_t items, uint32_t n_cu,
                       uint32_t per_cu, hipStream_t st, const FrameCacheDev* cache)
{
#define X(B, I) \
    if (block == B && items == I) return launch_cull_t<B, I>(job, s, n_cu, per_cu, st, cache);
    BZ_TUNING_TABLE(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_decide_only(const CullJob& job, const CompactScratch& s, uint32_t block, uint32_t items,
                              uint32_t n_cu, uint32_t per_cu, hipStream_t st, const InstHistDev* ih)
{
#define X(B, I)                                                                                                      \
    if (block == B && items == I)                                                                                     \
        return ih ? launch_decide_inst_t<B, I>(job, s, n_cu, per_cu, *ih, st)                                         \
                  : launch_cull_t<B, I>(job, s, n_cu, per_cu, st, nullptr, false);
    BZ_TUNING_TABLE(X)
#undef X
    return hipErrorInvalidValue;
}

void decide_inst_chunks(const CullJob& job, uint32_t block, uint32_t items, uint32_t n_cu, uint32_t per_cu,
                        uint32_t* tiles_per_chunk, uint32_t* n_long, uint32_t* n_chunks)
{
    uint32_t resident = 1;
#define X(B, I) \
    if (block == B && items == I) resident = inst_resident<B, I>();
    BZ_TUNING_TABLE(X)
#undef X
    // one chunk fewer than the workgroups that fit at once (one round, each workgroup a contiguous run of tiles, the
    // lengths differing by at most one tile): the scatter launch has a workgroup per chunk plus one for the commands,
    // and with as many chunks as slots it holds one workgroup more than three per CU (C1 size: scatter 15.6 us
    // against 12.8 us)
    uint64_t slots = (uint64_t)n_cu * (per_cu ? per_cu : resident);
    if (slots > 1) slots -= 1;
    if (slots > job.n_tiles) slots = job.n_tiles;
    if (slots == 0) slots = 1;
    *n_chunks = (uint32_t)slots;
    *tiles_per_chunk = (uint32_t)(job.n_tiles / slots);
    *n_long = (uint32_t)(job.n_tiles % slots);
    if (*tiles_per_chunk == 0) {  // (an empty job: one chunk of no tiles)
        *tiles_per_chunk = 1;
        *n_long = 0;
    }
}

hipError_t launch_late_cached(const CullJob& job_in, const CompactScratch& s, const FrameCacheDev& cache, uint32_t block,
                              uint32_t items, hipStream_t st)
{
    // the pass runs over super-tiles; the recorded tiles (block x items objects) are what the cache holds
    CullJob job = job_in;
    const uint32_t rec = block * items;
    const uint32_t n_rec_tiles = (job.n + rec - 1) / rec;
    job.n_tiles = late_cached_tiles(job.n);
    switch (rec) {
    case 512: return launch_late_cached_t<512>(job, s, cache, n_rec_tiles, st);
    case 1024: return launch_late_cached_t<1024>(job, s, cache, n_rec_tiles, st);
    case 2048: return launch_late_cached_t<2048>(job, s, cache, n_rec_tiles, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_cluster_objects(const CullJob& job, const CompactScratch& s, uint32_t* ccounts, uint32_t block,
                                  uint32_t items, hipStream_t st)
{
#define X(B, I) \
    if (block == B && items == I) return launch_cluster_objects_t<B, I>(job, s, ccounts, st);
    BZ_TUNING_TABLE(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_cull_batch(const CullJob* jobs_host, const CullJob* jobs_dev, const uint32_t* starts_dev,
                             uint32_t n_jobs, uint32_t total_tiles, uint32_t max_job_tiles, const CompactScratch& s,
                             uint32_t block, uint32_t items, hipStream_t st)
{
    if (n_jobs == 0 || total_tiles == 0) return hipSuccess;
#define X(B, I)                                                                                                      \
    if (block == B && items == I)                                                                                     \
        return launch_cull_batch_t<B, I>(jobs_host, jobs_dev, starts_dev, n_jobs, total_tiles, max_job_tiles, s, st);
    BZ_TUNING_TABLE(X)
#undef X
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------------------------------------
// exclusive prefix of tile counts (jobs of more than kScanTiles tiles)
// ------------------------------------------------------------------------------------------------
constexpr uint32_t kScanBlock = 1024;
constexpr uint32_t kScanPer = 8;  // consecutive counts per thread and round

// One workgroup walks the counts in rounds of kScanBlock * kScanPer. Each round is loaded and stored coalesced
// (thread t takes words t, t + kScanBlock, ...) through an LDS tile, in which thread t then owns kScanPer consecutive
// counts: an exclusive scan of its 8 counts, a wave scan of the thread totals and a block scan of the wave totals give
// each count its prefix (plus the rounds before); the last word is the total. The LDS tile has one spare word per 32,
// so the 8-word strided accesses spread over the banks.
__device__ __forceinline__ uint32_t scan_lds_index(uint32_t i) { return i + (i >> 5); }

__global__ __launch_bounds__(kScanBlock) void scan_counts_kernel(const uint32_t* __restrict__ counts, uint32_t n,
                                                                 uint32_t* __restrict__ prefix)
{
    constexpr uint32_t NW = kScanBlock / 64, R = kScanBlock * kScanPer;
    __shared__ uint32_t wsum[NW];
    __shared__ uint32_t tile[R + R / 32];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += R) {
        uint32_t c[kScanPer];
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t i = r0 + k * kScanBlock + tid;
            c[k] = i < n ? counts[i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) tile[scan_lds_index(k * kScanBlock + tid)] = c[k];
        __syncthreads();
        uint32_t v[kScanPer], own = 0;
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t x = tile[scan_lds_index(tid * kScanPer + k)];
            v[k] = own;
            own += x;
        }
        const uint32_t incl = wave_incl_scan(own, lane);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < NW; ++w) {
            before += w < wave ? wsum[w] : 0u;
            total += wsum[w];
        }
        const uint32_t base = carry + before + (incl - own);
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) tile[scan_lds_index(tid * kScanPer + k)] = base + v[k];
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kScanPer; ++k) {
            const uint32_t i = r0 + k * kScanBlock + tid;
            if (i < n) prefix[i] = tile[scan_lds_index(k * kScanBlock + tid)];
        }
        carry += total;
        __syncthreads();  // wsum and the tile are rewritten by the next round
    }
    if (threadIdx.x == 0) prefix[n] = carry;
}

hipError_t launch_scan_counts(const uint32_t* counts, uint32_t n, uint32_t* prefix, hipStream_t st)
{
    hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(kScanBlock), 0, st, counts, n, prefix);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t* p, uint32_t value, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, p, value, n);
    return hipGetLastError();
}

// Transform updates, in two launches so that an id named more than once in one commit takes its last occurrence
// (UpdateObjectTransform copies updates into the staging buffer in call order, vulkanDraw.cpp:816-820):
//   mark:    update i copies its id from the pinned slot into device memory and raises winner[id] to i + 1;
//   scatter: update i writes its 32-byte record into the position/scale and quaternion streams only if it is the
//            winner, which then clears winner[id] (the array is all zero between commits; a non-winner compares its
//            own i + 1 and never matches, whether it reads before or after that clear).
// Ids >= n_xf are skipped and set err[2].
__global__ void mark_staged_transforms_kernel(const uint32_t* __restrict__ ids, uint32_t n, uint32_t n_xf,
                                              uint32_t* dev_ids, uint32_t* winner, uint32_t* err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = ids[i];
    dev_ids[i] = id;
    if (id >= n_xf) {
        atomicOr(&err[2], 1u);
        return;
    }
    atomicMax(&winner[id], i + 1u);
}

__global__ void scatter_staged_transforms_kernel(const uint32_t* __restrict__ dev_ids,
                                                 const float4* __restrict__ records, uint32_t n, uint32_t n_xf,
                                                 uint32_t* winner, float4* xf_ps, float4* xf_q)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = dev_ids[i];
    if (id >= n_xf || winner[id] != i + 1u) return;
    const float4 ps = records[2u * i], q = records[2u * i + 1u];
    xf_ps[id] = ps;
    xf_q[id] = q;
    winner[id] = 0u;
}

hipError_t launch_scatter_staged_transforms(const uint32_t* ids, const float4* records, uint32_t n, uint32_t n_xf,
                                            uint32_t* dev_ids, uint32_t* winner, float4* xf_ps, float4* xf_q,
                                            uint32_t* err, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const dim3 g((n + 255) / 256), b(256);
    hipLaunchKernelGGL(mark_staged_transforms_kernel, g, b, 0, st, ids, n, n_xf, dev_ids, winner, err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scatter_staged_transforms_kernel, g, b, 0, st, (const uint32_t*)dev_ids, records, n, n_xf,
                       winner, xf_ps, xf_q);
    return hipGetLastError();
}

}  // namespace bz
