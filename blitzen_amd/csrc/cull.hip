// cull.hip — gfx950 draw-cull kernels (early / late / plain / temporal / onpc passes).
//
// Replaces InitialDrawCull.comp.glsl:12-58, LateDrawCull.comp.glsl:13-71, TransparentDrawCull.comp.glsl:13-65,
// OnpcDrawCull.comp.glsl:11-52 and drawOccTemporal.hlsl:13-64. The per-object arithmetic is in cull_math.h.
//
// Structure of one tile (BLOCK threads x ITEMS objects, objects tile*BLOCK*ITEMS + k*BLOCK + tid):
//   1. all ITEMS objects' streams are loaded first (coalesced SoA float4 / uint32 loads), then decided;
//   2. per item a wave64 ballot gives each emitting lane its rank inside the wave;
//   3. wave counts are scanned in LDS in (item, wave) order, which is ascending object order;
//   4. the tile's base in the output comes from a single-pass decoupled look-back over earlier tiles
//      (tiles are numbered by an atomic ticket in launch order, so every wait is on a running workgroup);
//   5. commands and visibility are written.
// The result order is ascending objectId, independent of scheduling (the reference uses atomicAdd order).
#include "bz_internal.h"

namespace bz {

namespace {

constexpr unsigned long long kFlagAgg = 1ull << 62;
constexpr unsigned long long kFlagInc = 2ull << 62;
constexpr unsigned long long kFlagMask = 3ull << 62;
constexpr uint32_t kSpinLimit = 1u << 24;  // bounded wait: ~seconds, then report an error and move on

__device__ __forceinline__ unsigned long long ld_status(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

// Decoupled look-back, executed by wave 0. `first_tile` is the first tile of this job in `status`.
// Returns the exclusive prefix (number of commands emitted by earlier tiles of the job).
__device__ __forceinline__ uint32_t lookback(unsigned long long* status, uint32_t tile, uint32_t first_tile,
                                             uint32_t agg, uint32_t lane, uint32_t* err)
{
    if (tile == first_tile) {
        if (lane == 0) st_status(&status[tile], kFlagInc | agg);
        return 0;
    }
    if (lane == 0) st_status(&status[tile], kFlagAgg | agg);
    uint32_t excl = 0;
    int64_t look = (int64_t)tile - 1;
    while (true) {
        const int64_t idx = look - (int64_t)lane;
        unsigned long long s = kFlagInc;  // before the job's first tile: inclusive 0
        if (idx >= (int64_t)first_tile) {
            uint32_t spins = 0;
            s = ld_status(&status[idx]);
            while ((s & kFlagMask) == 0) {
                if (++spins > kSpinLimit) {
                    atomicExch(err, 1u);
                    s = kFlagInc;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                s = ld_status(&status[idx]);
            }
        }
        const unsigned long long inc = __ballot((s & kFlagMask) == kFlagInc);
        const uint32_t val = (uint32_t)s;
        if (inc) {
            const uint32_t first = (uint32_t)__ffsll((long long)inc) - 1u;  // nearest tile with an inclusive value
            excl += wave_sum(lane <= first ? val : 0u);
            break;
        }
        excl += wave_sum(val);
        look -= 64;
    }
    if (lane == 0) st_status(&status[tile], kFlagInc | (excl + agg));
    return excl;
}

template <int BLOCK, int ITEMS>
struct TileSmem {
    uint32_t woff[ITEMS][BLOCK / 64];
    uint32_t prefix;
    uint32_t total;
    uint32_t tile;
    uint32_t job;
    uint32_t last;
};

// One tile of one job. Called by all threads of the workgroup.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void cull_tile(const CullJob& J, uint32_t tile_in_job, unsigned long long* status,
                                          uint32_t first_tile, uint32_t* err, TileSmem<BLOCK, ITEMS>& sm)
{
    constexpr int NW = BLOCK / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t pass = J.pass;
    const bool uses_vis = (pass == PASS_EARLY || pass == PASS_LATE);
    const SceneDev& S = J.scene;
    const uint32_t base = tile_in_job * (uint32_t)(BLOCK * ITEMS);

    uint32_t vis_prev[ITEMS], surf[ITEMS];
    float4 ps[ITEMS], q[ITEMS], sph[ITEMS];
    bool active[ITEMS];

    // ---- 1. loads
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = base + (uint32_t)k * BLOCK + tid;
        const bool valid = i < J.n;
        const uint32_t gi = J.begin + i;
        vis_prev[k] = (valid && uses_vis) ? J.vis_in[gi] : 0u;
        active[k] = valid && (pass != PASS_EARLY || vis_prev[k] != 0u);
        surf[k] = 0;
        ps[k] = q[k] = sph[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (active[k]) {
            const uint32_t x = S.obj_xf ? S.obj_xf[gi] : gi;
            surf[k] = S.obj_surf[gi];
            ps[k] = S.xf_ps[x];
            q[k] = S.xf_q[x];
        }
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
        if (active[k]) sph[k] = S.surf_sphere[surf[k]];

    // ---- 2. decisions
    bool emit[ITEMS], vis_new[ITEMS];
    uint2 draw[ITEMS];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        emit[k] = false;
        vis_new[k] = false;
        draw[k] = make_uint2(0u, 0u);
        if (active[k]) {
            float cx, cy, cz, r;
            sphere_to_view(J.view, sph[k].x, sph[k].y, sph[k].z, sph[k].w, ps[k].x, ps[k].y, ps[k].z, ps[k].w, q[k].x,
                           q[k].y, q[k].z, q[k].w, cx, cy, cz, r);
            bool visible = frustum_visible(J.view, cx, cy, cz, r);
            if (J.hiz && visible) {
                float a0, a1, a2, a3;
                if (project_sphere(cx, cy, cz, r, J.view.znear, J.view.p00, J.view.p11, a0, a1, a2, a3))
                    visible = occlusion_visible(J.view, J.pyr, a0, a1, a2, a3, cz, r);
            }
            vis_new[k] = visible;
            emit[k] = (pass == PASS_LATE) ? (visible && vis_prev[k] == 0u) : visible;
            if (emit[k]) {
                const uint2 lo = S.surf_lod[surf[k]];
                const float len = sqrtf((cx * cx + cy * cy) + cz * cz);
                const float distance = glsl_max(len - r, 0.0f);
                const float threshold = (distance * J.view.lodTarget) / ps[k].w;
                uint32_t idx = 0;
                for (uint32_t l = 1; l < lo.y; ++l)
                    if (S.lod_error[lo.x + l] < threshold) idx = l;
                const uint32_t abs_idx = (pass == PASS_ONPC) ? idx : idx + lo.x;
                draw[k] = S.lod_draw[abs_idx];
            }
        }
    }

    // ---- 3. wave ballots + LDS scan in (item, wave) order
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const unsigned long long m = __ballot(emit[k]);
        rank[k] = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) sm.woff[k][wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (wave == 0) {
        constexpr uint32_t E = ITEMS * NW;  // <= 64
        const uint32_t c = lane < E ? (&sm.woff[0][0])[lane] : 0u;
        const uint32_t incl = wave_incl_scan(c, lane);
        const uint32_t agg = __shfl(incl, 63, 64);
        if (lane < E) (&sm.woff[0][0])[lane] = incl - c;
        const uint32_t prefix = lookback(status, J.tile_start + tile_in_job, first_tile, agg, lane, err);
        if (lane == 0) {
            sm.prefix = prefix;
            sm.total = prefix + agg;
        }
    }
    __syncthreads();

    // ---- 4. outputs
    const uint32_t prefix = sm.prefix;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = base + (uint32_t)k * BLOCK + tid;
        if (emit[k]) {
            const uint32_t pos = prefix + sm.woff[k][wave] + rank[k];
            if (pos < J.capacity) {
                uint2* c = reinterpret_cast<uint2*>(J.cmds + (size_t)pos * 6u);
                c[0] = make_uint2(J.begin + i, draw[k].x);  // objectId, indexCount
                c[1] = make_uint2(1u, draw[k].y);           // instanceCount, firstIndex
                c[2] = make_uint2(0u, 0u);                  // vertexOffset, firstInstance
            }
        }
        if (pass == PASS_LATE && i < J.n) J.vis_out[J.begin + i] = vis_new[k] ? 1u : 0u;
    }
    if (tid == 0 && tile_in_job + 1 == J.n_tiles) *J.count = sm.total;
}

// Last workgroup out clears the look-back state for the next launch. No memory fence is needed: every status
// access is an agent-scope atomic, and a workgroup's look-back reads have returned before it counts itself done.
// (An agent-scope release fence would write back the whole L2 of the XCD once per workgroup.)
template <int BLOCK>
__device__ __forceinline__ void finish_and_reset(const LookbackScratch& s, uint32_t n_status, uint32_t n_wg,
                                                 uint32_t* sm_last)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t d = atomicAdd(s.done, 1u);
        *sm_last = (d + 1u == n_wg) ? 1u : 0u;
    }
    __syncthreads();
    if (*sm_last) {
        for (uint32_t t = threadIdx.x; t < n_status; t += BLOCK) st_status(&s.status[t], 0ull);
        if (threadIdx.x == 0) {
            atomicExch(s.ticket, 0u);
            atomicExch(s.done, 0u);
        }
    }
}

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void cull_kernel(CullJob J, LookbackScratch s)
{
    __shared__ TileSmem<BLOCK, ITEMS> sm;
    if (threadIdx.x == 0) sm.tile = atomicAdd(s.ticket, 1u);
    __syncthreads();
    const uint32_t tile = sm.tile;
    cull_tile<BLOCK, ITEMS>(J, tile, s.status, 0u, s.error, sm);
    finish_and_reset<BLOCK>(s, J.n_tiles, gridDim.x, &sm.last);
}

// Persistent batched kernel: workgroups take (job, tile) items from a global ticket until all are done.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void cull_batch_kernel(const CullJob* __restrict__ jobs, uint32_t n_jobs,
                                                           uint32_t total_tiles, LookbackScratch s)
{
    __shared__ TileSmem<BLOCK, ITEMS> sm;
    __shared__ CullJob sjob;
    uint32_t cached_job = 0xffffffffu;
    while (true) {
        if (threadIdx.x == 0) {
            const uint32_t t = atomicAdd(s.ticket, 1u);
            sm.tile = t;
            if (t < total_tiles) {
                // binary search: last job with tile_start <= t
                uint32_t lo = 0, hi = n_jobs - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (jobs[mid].tile_start <= t) lo = mid;
                    else hi = mid - 1;
                }
                sm.job = lo;
            }
        }
        __syncthreads();
        const uint32_t t = sm.tile;
        if (t >= total_tiles) break;
        const uint32_t j = sm.job;
        if (j != cached_job) {
            // stage the job descriptor in LDS (dword copy)
            const uint32_t* src = reinterpret_cast<const uint32_t*>(&jobs[j]);
            uint32_t* dst = reinterpret_cast<uint32_t*>(&sjob);
            for (uint32_t w = threadIdx.x; w < sizeof(CullJob) / 4; w += BLOCK) dst[w] = src[w];
            cached_job = j;
            __syncthreads();
        }
        cull_tile<BLOCK, ITEMS>(sjob, t - sjob.tile_start, s.status, sjob.tile_start, s.error, sm);
        __syncthreads();
    }
    finish_and_reset<BLOCK>(s, total_tiles, gridDim.x, &sm.last);
}

template <int BLOCK, int ITEMS>
hipError_t launch_cull_t(const CullJob& job, const LookbackScratch& s, hipStream_t st)
{
    hipLaunchKernelGGL((cull_kernel<BLOCK, ITEMS>), dim3(job.n_tiles), dim3(BLOCK), 0, st, job, s);
    return hipGetLastError();
}

template <int BLOCK, int ITEMS>
hipError_t launch_cull_batch_t(const CullJob* jobs, uint32_t n_jobs, uint32_t total_tiles, const LookbackScratch& s,
                               uint32_t grid, hipStream_t st)
{
    hipLaunchKernelGGL((cull_batch_kernel<BLOCK, ITEMS>), dim3(grid), dim3(BLOCK), 0, st, jobs, n_jobs, total_tiles, s);
    return hipGetLastError();
}

}  // namespace

#define BZ_TUNING_TABLE(X) \
    X(64, 1) X(64, 2) X(64, 4) X(64, 8) \
    X(128, 1) X(128, 2) X(128, 4) X(128, 8) \
    X(256, 1) X(256, 2) X(256, 4) X(256, 8) \
    X(512, 1) X(512, 2) X(512, 4) X(512, 8)

hipError_t launch_cull(const CullJob& job, const LookbackScratch& s, uint32_t block, uint32_t items, hipStream_t st)
{
#define X(B, I) \
    if (block == B && items == I) return launch_cull_t<B, I>(job, s, st);
    BZ_TUNING_TABLE(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_cull_batch(const CullJob* jobs_dev, uint32_t n_jobs, uint32_t total_tiles, const LookbackScratch& s,
                             uint32_t block, uint32_t items, uint32_t grid, hipStream_t st)
{
#define X(B, I) \
    if (block == B && items == I) return launch_cull_batch_t<B, I>(jobs_dev, n_jobs, total_tiles, s, grid, st);
    BZ_TUNING_TABLE(X)
#undef X
    return hipErrorInvalidValue;
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

__global__ void scatter_transforms_kernel(const uint32_t* __restrict__ ids, const float4* __restrict__ src_ps,
                                          const float4* __restrict__ src_q, uint32_t n, float4* xf_ps, float4* xf_q)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = ids[i];
    xf_ps[id] = src_ps[i];
    xf_q[id] = src_q[i];
}

hipError_t launch_scatter_transforms(const uint32_t* ids, const float4* src_ps, const float4* src_q, uint32_t n,
                                     float4* xf_ps, float4* xf_q, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_transforms_kernel, dim3((n + 255) / 256), dim3(256), 0, st, ids, src_ps, src_q, n, xf_ps,
                       xf_q);
    return hipGetLastError();
}

}  // namespace bz
