// pyramid.hip — min-depth (Hi-Z) pyramid for gfx950, and the device depth rasteriser for synthetic requests.
//
// Replaces DepthPyramidGeneration.comp.glsl:13-19 dispatched once per mip with a barrier in between
// (vulkanDraw.cpp:554-622). Each level texel is the MIN over the non-zero-weight texels of the bilinear
// footprint of its centre in the level above (level 0 samples the depth buffer), the rule in cull_math.h.
//
// Single pass for power-of-two bases: each workgroup owns a tile of level 0 (up to 32 x 32 texels), writes it,
// and reduces it in LDS down to 1 x 1 (levels 1..5, exact 2:1 with edge clamp). The last workgroup
// (the highest-numbered one, waiting on flagged words) reduces the remaining levels from the
// 1-texel-per-tile image. Non-power-of-two
// bases fall back to one launch per level with the general footprint rule.
#include "bz_internal.h"
#include "synth_rng.h"

namespace bz {

namespace {

constexpr uint32_t kPyrBlock = 256;
constexpr uint32_t kTile = 32;
constexpr uint32_t kTailMax = 4096;  // texels of the per-tile image the last workgroup reduces in LDS
constexpr unsigned long long kTailFlag = 1ull << 63;
constexpr uint32_t kTailSpinLimit = 1u << 20;

__device__ __forceinline__ float min4(float a, float b, float c, float d) { return glsl_min(glsl_min(a, b), glsl_min(c, d)); }

// Reduce src (sw x sh, LDS) to dst (dw x dh = max(1, s/2)) exactly 2:1 with clamp; optionally write to global.
__device__ __forceinline__ void reduce_lds(const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw,
                                           uint32_t dh, float* gdst, uint32_t gpitch, uint32_t gx, uint32_t gy)
{
    for (uint32_t t = threadIdx.x; t < dw * dh; t += kPyrBlock) {
        const uint32_t lx = t % dw, ly = t / dw;
        const uint32_t x0 = 2u * lx, y0 = 2u * ly;
        const uint32_t x1 = sw > 1u ? x0 + 1u : x0;
        const uint32_t y1 = sh > 1u ? y0 + 1u : y0;
        const float d = min4(src[y0 * sw + x0], src[y0 * sw + x1], src[y1 * sw + x0], src[y1 * sw + x1]);
        dst[t] = d;
        gdst[(size_t)(gy + ly) * gpitch + gx + lx] = d;
    }
}

__device__ void pyramid_tile(const PyrBuildJob& J, uint32_t tile, float* bufA, float* bufB, uint32_t* s_last,
                             uint32_t* err)
{
    const uint32_t tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    const uint32_t tw = J.tile_w, th = J.tile_h;
    // level 0 from the depth buffer (general footprint: the depth buffer need not be 2:1 of the base)
    const float fbw = (float)J.base_w, fbh = (float)J.base_h;
    for (uint32_t t = threadIdx.x; t < tw * th; t += kPyrBlock) {
        const uint32_t lx = t % tw, ly = t / tw;
        const uint32_t x = tx * tw + lx, y = ty * th + ly;
        const float u = ((float)x + 0.5f) / fbw;
        const float v = ((float)y + 0.5f) / fbh;
        const Footprint f = footprint(u, v, J.w, J.h);
        const float d = footprint_min(J.depth, J.w, f);
        bufA[t] = d;
        J.texels[J.level_offset[0] + (size_t)y * J.base_w + x] = d;
    }
    __syncthreads();
    uint32_t cw = tw, ch = th;
    float* src = bufA;
    float* dst = bufB;
    for (uint32_t L = 1; L < J.tile_levels; ++L) {
        const uint32_t nw = cw > 1u ? cw >> 1 : 1u, nh = ch > 1u ? ch >> 1 : 1u;
        reduce_lds(src, cw, ch, dst, nw, nh, J.texels + J.level_offset[L], level_dim(J.base_w, L), tx * nw, ty * nh);
        __syncthreads();
        float* tmp = src;
        src = dst;
        dst = tmp;
        cw = nw;
        ch = nh;
    }
    if (J.tile_levels >= J.mips) return;
    // ---- hand each tile's 1x1 value to the highest-numbered workgroup, which reduces the tail levels. The value
    // travels in a 64-bit agent-scope atomic word with a presence flag; the receiving workgroup clears each word
    // after reading it, so every build starts with all words zero. Neither a release fence (an L2 write-back per
    // workgroup) nor a shared completion counter (a serialised atomic per workgroup) is needed.
    const uint32_t n_tiles = J.tiles_x * J.tiles_y;
    if (tile + 1u < n_tiles) {
        if (threadIdx.x == 0)
            __hip_atomic_store(&J.tail[tile], kTailFlag | __float_as_uint(src[0]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    (void)s_last;
    const uint32_t L0 = J.tile_levels - 1u;
    cw = level_dim(J.base_w, L0);
    ch = level_dim(J.base_h, L0);
    const float own = src[0];
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < cw * ch; t += kPyrBlock) {
        if (t == tile) {
            bufA[t] = own;
            continue;
        }
        unsigned long long w = __hip_atomic_load(&J.tail[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spins = 0;
        while (!(w & kTailFlag)) {
            if (++spins > kTailSpinLimit) {
                atomicExch(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            w = __hip_atomic_load(&J.tail[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        bufA[t] = __uint_as_float((uint32_t)w);
        __hip_atomic_store(&J.tail[t], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    src = bufA;
    dst = bufB;
    for (uint32_t L = L0 + 1u; L < J.mips; ++L) {
        const uint32_t nw = cw > 1u ? cw >> 1 : 1u, nh = ch > 1u ? ch >> 1 : 1u;
        reduce_lds(src, cw, ch, dst, nw, nh, J.texels + J.level_offset[L], nw, 0u, 0u);
        __syncthreads();
        float* tmp = src;
        src = dst;
        dst = tmp;
        cw = nw;
        ch = nh;
    }
}

__global__ __launch_bounds__(kPyrBlock) void pyramid_kernel(PyrBuildJob J, uint32_t* err)
{
    __shared__ float bufA[kTailMax];
    __shared__ float bufB[kTailMax / 2];
    __shared__ uint32_t s_last;
    pyramid_tile(J, blockIdx.x, bufA, bufB, &s_last, err);
}

__global__ __launch_bounds__(kPyrBlock) void pyramid_batch_kernel(const PyrBuildJob* __restrict__ jobs, uint32_t n_jobs,
                                                                  uint32_t* err)
{
    __shared__ float bufA[kTailMax];
    __shared__ float bufB[kTailMax / 2];
    __shared__ uint32_t s_last;
    __shared__ uint32_t s_job;
    if (threadIdx.x == 0) {
        uint32_t lo = 0, hi = n_jobs - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (jobs[mid].wg_start <= blockIdx.x) lo = mid;
            else hi = mid - 1;
        }
        s_job = lo;
    }
    __syncthreads();
    const PyrBuildJob J = jobs[s_job];
    pyramid_tile(J, blockIdx.x - J.wg_start, bufA, bufB, &s_last, err);
}

// General per-level footprint reduction (any sizes).
__global__ void pyramid_level_kernel(const float* __restrict__ src, uint32_t sw, uint32_t sh, float* __restrict__ dst,
                                     uint32_t dw, uint32_t dh)
{
    const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (x >= dw || y >= dh) return;
    const float u = ((float)x + 0.5f) / (float)dw;
    const float v = ((float)y + 0.5f) / (float)dh;
    const Footprint f = footprint(u, v, sw, sh);
    dst[(size_t)y * dw + x] = footprint_min(src, sw, f);
}

// Synthetic depth: one workgroup per 64 x 64 pixel tile. The workgroup regenerates every rectangle from the seed,
// keeps those overlapping its tile in LDS, then each pixel takes the max (nearest, reverse-z) of the ground plane
// and its covering rectangles. max is exact, so the order of the LDS list does not matter.
constexpr uint32_t kRasterTile = 64;
constexpr uint32_t kMaxRects = 4096;

__global__ __launch_bounds__(256) void synth_depth_kernel(uint64_t seed, uint32_t W, uint32_t H, uint32_t n_rects,
                                                          float p11, float znear, float* __restrict__ depth)
{
    __shared__ int4 s_rect[kMaxRects];
    __shared__ float s_depth[kMaxRects];
    __shared__ uint32_t s_n;
    const int32_t tx0 = (int32_t)(blockIdx.x * kRasterTile), ty0 = (int32_t)(blockIdx.y * kRasterTile);
    const int32_t tx1 = tx0 + (int32_t)kRasterTile, ty1 = ty0 + (int32_t)kRasterTile;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < n_rects; r += 256) {
        const SynthRect s = synth_rect(seed, W, H, r);
        if (s.x0 < tx1 && s.x1 > tx0 && s.y0 < ty1 && s.y1 > ty0) {
            const uint32_t slot = atomicAdd(&s_n, 1u);
            s_rect[slot] = make_int4(s.x0, s.y0, s.x1, s.y1);
            s_depth[slot] = s.depth;
        }
    }
    __syncthreads();
    const uint32_t n = s_n;
    for (uint32_t t = threadIdx.x; t < kRasterTile * kRasterTile; t += 256) {
        const int32_t x = tx0 + (int32_t)(t % kRasterTile), y = ty0 + (int32_t)(t / kRasterTile);
        if (x >= (int32_t)W || y >= (int32_t)H) continue;
        float d = synth_plane_depth((uint32_t)y, H, p11, znear);
        for (uint32_t k = 0; k < n; ++k) {
            const int4 rc = s_rect[k];
            if (x >= rc.x && x < rc.z && y >= rc.y && y < rc.w) d = glsl_max(d, s_depth[k]);
        }
        depth[(size_t)y * W + x] = d;
    }
}

}  // namespace

hipError_t launch_pyramid(const PyrBuildJob& job, uint32_t* err, hipStream_t st)
{
    hipLaunchKernelGGL(pyramid_kernel, dim3(job.tiles_x * job.tiles_y), dim3(kPyrBlock), 0, st, job, err);
    return hipGetLastError();
}

hipError_t launch_pyramid_batch(const PyrBuildJob* jobs_dev, uint32_t n_jobs, uint32_t total_wgs, uint32_t max_mips,
                                uint32_t* err, hipStream_t st)
{
    (void)max_mips;
    if (n_jobs == 0 || total_wgs == 0) return hipSuccess;
    hipLaunchKernelGGL(pyramid_batch_kernel, dim3(total_wgs), dim3(kPyrBlock), 0, st, jobs_dev, n_jobs, err);
    return hipGetLastError();
}

hipError_t launch_pyramid_level(const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw, uint32_t dh,
                                hipStream_t st)
{
    hipLaunchKernelGGL(pyramid_level_kernel, dim3((dw + 15) / 16, (dh + 15) / 16), dim3(256), 0, st, src, sw, sh, dst,
                       dw, dh);
    return hipGetLastError();
}

hipError_t launch_synth_depth(uint64_t seed, uint32_t w, uint32_t h, uint32_t occluders, float p11, float znear,
                              float* depth, hipStream_t st)
{
    if (occluders > kMaxRects) return hipErrorInvalidValue;
    hipLaunchKernelGGL(synth_depth_kernel, dim3((w + kRasterTile - 1) / kRasterTile, (h + kRasterTile - 1) / kRasterTile),
                       dim3(256), 0, st, seed, w, h, occluders, p11, znear, depth);
    return hipGetLastError();
}

}  // namespace bz
