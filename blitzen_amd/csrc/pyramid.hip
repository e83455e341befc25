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
constexpr unsigned long long kTailFlag = 1ull << 63;
constexpr uint32_t kLdsJobs = 1024;  // batched: job start table staged in LDS up to this many jobs

__device__ __forceinline__ float min4(float a, float b, float c, float d) { return glsl_min(glsl_min(a, b), glsl_min(c, d)); }

// Reduce src (sw x sh, LDS) to dst (dw x dh = max(1, s/2)) exactly 2:1 with clamp, and write it to global.
// All extents are powers of two (single-pass pyramids only).
template <uint32_t NT>
__device__ __forceinline__ void reduce_lds(const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw,
                                           uint32_t dh, float* gdst, uint32_t gpitch, uint32_t gx, uint32_t gy)
{
    const uint32_t dshift = (uint32_t)__builtin_ctz(dw);
    for (uint32_t t = threadIdx.x; t < dw * dh; t += NT) {
        const uint32_t lx = t & (dw - 1u), ly = t >> dshift;
        const uint32_t x0 = 2u * lx, y0 = 2u * ly;
        const uint32_t x1 = sw > 1u ? x0 + 1u : x0;
        const uint32_t y1 = sh > 1u ? y0 + 1u : y0;
        const float d = min4(src[y0 * sw + x0], src[y0 * sw + x1], src[y1 * sw + x0], src[y1 * sw + x1]);
        dst[t] = d;
        gdst[(size_t)(gy + ly) * gpitch + gx + lx] = d;
    }
}

// Waits for one flagged per-tile word of the tail, already loaded once into w (the wait is bounded; giving up sets
// *err), and clears it after reading, so the next build starts from zero.
__device__ __forceinline__ float tail_finish(const PyrBuildJob& J, uint32_t idx, unsigned long long w, uint32_t* err)
{
    uint32_t spins = 0;
    while (!(w & kTailFlag)) {
        if (++spins > J.spin_limit) {
            atomicExch(err, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        w = __hip_atomic_load(&J.tail[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_store(&J.tail[idx], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __uint_as_float((uint32_t)w);
}

// Reads one per-tile word for the tail: the owner's own value, or the flagged word another workgroup stored
// (cleared after reading, so the next build starts from zero). The wait is bounded; giving up sets *err.
__device__ __forceinline__ float tail_word(const PyrBuildJob& J, uint32_t idx, uint32_t own_tile, float own,
                                           uint32_t* err)
{
    if (idx == own_tile) return own;
    unsigned long long w = __hip_atomic_load(&J.tail[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (!(w & kTailFlag)) {
        if (++spins > J.spin_limit) {
            atomicExch(err, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        w = __hip_atomic_load(&J.tail[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_store(&J.tail[idx], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __uint_as_float((uint32_t)w);
}

constexpr uint32_t kTileTexels = kTile * kTile;
constexpr uint32_t kPyrBlockWide = 512;  // single-view builds whose level 0 uses the general footprint rule

// One workgroup's tile of a single-pass (power-of-two) pyramid. bufA holds kTileTexels floats, bufB half that.
// NT threads per workgroup: 256 for exact 2:1 level 0 (and batches), 512 for the general footprint rule of a single
// build (more gathers in flight per tile: measured 5% faster there, slower for 2:1).
template <uint32_t NT>
__device__ void pyramid_tile(const PyrBuildJob& J, uint32_t tile, float* bufA, float* bufB, uint32_t* err)
{
    constexpr uint32_t kPerThread = kTileTexels / NT;
    const uint32_t tid = threadIdx.x;
    const uint32_t tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    const uint32_t tw = J.tile_w, th = J.tile_h, n0 = tw * th;
    const uint32_t tshift = (uint32_t)__builtin_ctz(tw);
    float* lvl0 = J.texels + J.level_offset[0];
    // ---- level 0 from the depth buffer; every load of a thread is issued before the first min
    if (J.w == 2u * J.base_w && J.h == 2u * J.base_h && ((uintptr_t)J.depth & 7u) == 0) {
        // exactly 2:1: the footprint of texel (x, y) is depth {2x, 2x+1} x {2y, 2y+1} (frac(tx) = 0.5 on both axes)
        float2 r0[kPerThread], r1[kPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t t0 = tid + k * NT, t = t0 < n0 ? t0 : 0u;
            const uint32_t x = tx * tw + (t & (tw - 1u)), y = ty * th + (t >> tshift);
            const float2* row = reinterpret_cast<const float2*>(J.depth + (size_t)(2u * y) * J.w) + x;
            r0[k] = row[0];
            r1[k] = row[J.w / 2u];
        }
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t t = tid + k * NT;
            if (t < n0) {
                const uint32_t x = tx * tw + (t & (tw - 1u)), y = ty * th + (t >> tshift);
                const float d = min4(r0[k].x, r0[k].y, r1[k].x, r1[k].y);
                bufA[t] = d;
                lvl0[(size_t)y * J.base_w + x] = d;
            }
        }
    } else {
        // general ratio: the footprint rule; the base extent is a power of two, so (x + 0.5) / W equals
        // (x + 0.5) * (1 / W) exactly
        const float rbw = 1.0f / (float)J.base_w, rbh = 1.0f / (float)J.base_h;
        float d00[kPerThread], d01[kPerThread], d10[kPerThread], d11[kPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t t0 = tid + k * NT, t = t0 < n0 ? t0 : 0u;
            const uint32_t x = tx * tw + (t & (tw - 1u)), y = ty * th + (t >> tshift);
            const Footprint f = footprint(((float)x + 0.5f) * rbw, ((float)y + 0.5f) * rbh, J.w, J.h);
            d00[k] = J.depth[(size_t)f.y0 * J.w + f.x0];
            d01[k] = J.depth[(size_t)f.y0 * J.w + f.x1];
            d10[k] = J.depth[(size_t)f.y1 * J.w + f.x0];
            d11[k] = J.depth[(size_t)f.y1 * J.w + f.x1];
        }
#pragma unroll
        for (uint32_t k = 0; k < kPerThread; ++k) {
            const uint32_t t = tid + k * NT;
            if (t < n0) {
                const uint32_t x = tx * tw + (t & (tw - 1u)), y = ty * th + (t >> tshift);
                const float d = min4(d00[k], d01[k], d10[k], d11[k]);
                bufA[t] = d;
                lvl0[(size_t)y * J.base_w + x] = d;
            }
        }
    }
    __syncthreads();
    // ---- levels inside the tile, in LDS
    uint32_t cw = tw, ch = th;
    float* src = bufA;
    float* dst = bufB;
    for (uint32_t L = 1; L < J.tile_levels; ++L) {
        const uint32_t nw = cw > 1u ? cw >> 1 : 1u, nh = ch > 1u ? ch >> 1 : 1u;
        reduce_lds<NT>(src, cw, ch, dst, nw, nh, J.texels + J.level_offset[L], level_dim(J.base_w, L), tx * nw, ty * nh);
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
        if (tid == 0)
            __hip_atomic_store(&J.tail[tile], kTailFlag | __float_as_uint(src[0]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const float own = src[0];
    __syncthreads();  // every thread has read `own` before bufA is rewritten
    const uint32_t L0 = J.tile_levels - 1u;
    cw = level_dim(J.base_w, L0);
    ch = level_dim(J.base_h, L0);
    if (n_tiles <= kTileTexels) {
        // the whole tile image (one texel per tile, level L0) fits the LDS tile: every word of a thread is loaded
        // before the first wait (one round trip), then the remaining levels are reduced in LDS
        constexpr uint32_t KW = kTileTexels / NT;
        unsigned long long w[KW];
#pragma unroll
        for (uint32_t k = 0; k < KW; ++k) {
            const uint32_t i = tid + k * NT;
            w[k] = (i < n_tiles && i != tile)
                       ? __hip_atomic_load(&J.tail[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                       : 0ull;
        }
#pragma unroll
        for (uint32_t k = 0; k < KW; ++k) {
            const uint32_t i = tid + k * NT;
            if (i < n_tiles) bufA[i] = i == tile ? own : tail_finish(J, i, w[k], err);
        }
        __syncthreads();
        src = bufA;
        dst = bufB;
        for (uint32_t L = L0 + 1u; L < J.mips; ++L) {
            const uint32_t nw = cw > 1u ? cw >> 1 : 1u, nh = ch > 1u ? ch >> 1 : 1u;
            reduce_lds<NT>(src, cw, ch, dst, nw, nh, J.texels + J.level_offset[L], nw, 0u, 0u);
            __syncthreads();
            float* tmp = src;
            src = dst;
            dst = tmp;
            cw = nw;
            ch = nh;
        }
        return;
    }
    // larger tile images: reduced to the next level while they are collected
    {
        const uint32_t nw = cw > 1u ? cw >> 1 : 1u, nh = ch > 1u ? ch >> 1 : 1u;
        const uint32_t nshift = (uint32_t)__builtin_ctz(nw);
        float* g = J.texels + J.level_offset[L0 + 1u];
        for (uint32_t t = tid; t < nw * nh; t += NT) {
            const uint32_t x0 = 2u * (t & (nw - 1u)), y0 = 2u * (t >> nshift);
            const bool hx = cw > 1u, hy = ch > 1u;  // a second column / row exists (else the clamp repeats the first)
            const float v00 = tail_word(J, y0 * cw + x0, tile, own, err);
            const float v01 = hx ? tail_word(J, y0 * cw + x0 + 1u, tile, own, err) : v00;
            const float v10 = hy ? tail_word(J, (y0 + 1u) * cw + x0, tile, own, err) : v00;
            const float v11 = (hx && hy) ? tail_word(J, (y0 + 1u) * cw + x0 + 1u, tile, own, err) : (hx ? v01 : v10);
            const float d = min4(v00, v01, v10, v11);
            bufA[t] = d;
            g[t] = d;
        }
        cw = nw;
        ch = nh;
    }
    __syncthreads();
    src = bufA;
    dst = bufB;
    for (uint32_t L = L0 + 2u; L < J.mips; ++L) {
        const uint32_t nw = cw > 1u ? cw >> 1 : 1u, nh = ch > 1u ? ch >> 1 : 1u;
        reduce_lds<NT>(src, cw, ch, dst, nw, nh, J.texels + J.level_offset[L], nw, 0u, 0u);
        __syncthreads();
        float* tmp = src;
        src = dst;
        dst = tmp;
        cw = nw;
        ch = nh;
    }
}

template <uint32_t NT>
__global__ __launch_bounds__(NT) void pyramid_kernel(PyrBuildJob J, uint32_t* err)
{
    __shared__ float bufA[kTileTexels];
    __shared__ float bufB[kTileTexels / 2];
    pyramid_tile<NT>(J, blockIdx.x, bufA, bufB, err);
}

// Last entry of the ascending table `starts` that is <= x (searched in LDS when staged there).
__device__ __forceinline__ uint32_t last_le(const uint32_t* starts, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (starts[mid] <= x) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

// Many pyramids in one launch: workgroup (x, y) builds level-0 tile x of job y; workgroups past the job's tiles return
// at once. The job index is the workgroup's y, so no table of first workgroups is searched (the search's load cost
// each workgroup a round trip: the 256-view batch 336.6 us against 314.6 us); the descriptor is copied to LDS. Within
// a job, workgroups are dispatched in tile order, so the job's last tile (which waits on the others' hand-off words)
// starts after every tile it waits for.
__global__ __launch_bounds__(kPyrBlock) void pyramid_batch_kernel(const PyrBuildJob* __restrict__ jobs, uint32_t* err)
{
    __shared__ float bufA[kTileTexels];
    __shared__ float bufB[kTileTexels / 2];
    __shared__ PyrBuildJob sJ;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&jobs[blockIdx.y]);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&sJ);
    for (uint32_t w = threadIdx.x; w < sizeof(PyrBuildJob) / 4; w += kPyrBlock) dst[w] = src[w];
    __syncthreads();
    if (blockIdx.x >= sJ.tiles_x * sJ.tiles_y) return;
    pyramid_tile<kPyrBlock>(sJ, blockIdx.x, bufA, bufB, err);
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

// Synthetic depth: one workgroup per 64 x 64 pixel tile. The workgroup regenerates the rectangles from the seed in
// chunks, keeps those overlapping its tile in LDS, and each pixel (16 per thread, in registers) takes the max
// (nearest, reverse-z) of the ground plane and its covering rectangles. max is exact, so the order of the LDS list
// does not matter.
constexpr uint32_t kRasterTile = kSynthTile;
constexpr uint32_t kMaxRects = kSynthMaxRects;
constexpr uint32_t kRectChunk = 512;
constexpr uint32_t kRasterPix = kRasterTile * kRasterTile / 256;

__device__ void synth_depth_tile(uint64_t seed, uint32_t W, uint32_t H, uint32_t n_rects, float p11, float znear,
                                 float* __restrict__ depth, uint32_t bx, uint32_t by, int4* s_rect, float* s_depth,
                                 uint32_t* s_n)
{
    const int32_t tx0 = (int32_t)(bx * kRasterTile), ty0 = (int32_t)(by * kRasterTile);
    const int32_t tx1 = tx0 + (int32_t)kRasterTile, ty1 = ty0 + (int32_t)kRasterTile;
    float d[kRasterPix];
#pragma unroll
    for (uint32_t k = 0; k < kRasterPix; ++k) {
        const uint32_t t = threadIdx.x + k * 256u;
        const int32_t y = ty0 + (int32_t)(t / kRasterTile);
        d[k] = synth_plane_depth((uint32_t)(y < (int32_t)H ? y : 0), H, p11, znear);
    }
    for (uint32_t r0 = 0; r0 < n_rects; r0 += kRectChunk) {
        if (threadIdx.x == 0) *s_n = 0;
        __syncthreads();
        const uint32_t r1 = r0 + kRectChunk < n_rects ? r0 + kRectChunk : n_rects;
        for (uint32_t r = r0 + threadIdx.x; r < r1; r += 256) {
            const SynthRect sr = synth_rect(seed, W, H, r);
            if (sr.x0 < tx1 && sr.x1 > tx0 && sr.y0 < ty1 && sr.y1 > ty0) {
                const uint32_t slot = atomicAdd(s_n, 1u);
                s_rect[slot] = make_int4(sr.x0, sr.y0, sr.x1, sr.y1);
                s_depth[slot] = sr.depth;
            }
        }
        __syncthreads();
        const uint32_t n = *s_n;
        for (uint32_t j = 0; j < n; ++j) {
            const int4 rc = s_rect[j];
            const float rd = s_depth[j];
#pragma unroll
            for (uint32_t k = 0; k < kRasterPix; ++k) {
                const uint32_t t = threadIdx.x + k * 256u;
                const int32_t x = tx0 + (int32_t)(t % kRasterTile), y = ty0 + (int32_t)(t / kRasterTile);
                if (x >= rc.x && x < rc.z && y >= rc.y && y < rc.w) d[k] = glsl_max(d[k], rd);
            }
        }
        __syncthreads();  // the list is rewritten by the next chunk
    }
#pragma unroll
    for (uint32_t k = 0; k < kRasterPix; ++k) {
        const uint32_t t = threadIdx.x + k * 256u;
        const int32_t x = tx0 + (int32_t)(t % kRasterTile), y = ty0 + (int32_t)(t / kRasterTile);
        if (x < (int32_t)W && y < (int32_t)H) depth[(size_t)y * W + x] = d[k];
    }
}

__global__ __launch_bounds__(256) void synth_depth_kernel(uint64_t seed, uint32_t W, uint32_t H, uint32_t n_rects,
                                                          float p11, float znear, float* __restrict__ depth)
{
    __shared__ int4 s_rect[kRectChunk];
    __shared__ float s_depth[kRectChunk];
    __shared__ uint32_t s_n;
    synth_depth_tile(seed, W, H, n_rects, p11, znear, depth, blockIdx.x, blockIdx.y, s_rect, s_depth, &s_n);
}

// Many depth buffers in one launch (workgroup b rasterises tile b - wg_start of the job whose range holds b).
__global__ __launch_bounds__(256) void synth_depth_batch_kernel(const SynthJob* __restrict__ jobs,
                                                                const uint32_t* __restrict__ starts, uint32_t n_jobs)
{
    __shared__ int4 s_rect[kRectChunk];
    __shared__ float s_depth[kRectChunk];
    __shared__ uint32_t s_n;
    __shared__ uint32_t s_start[kLdsJobs];
    uint32_t j;
    if (n_jobs <= kLdsJobs) {
        for (uint32_t k = threadIdx.x; k < n_jobs; k += 256) s_start[k] = starts[k];
        __syncthreads();
        j = last_le(s_start, n_jobs, blockIdx.x);
    } else {
        j = last_le(starts, n_jobs, blockIdx.x);
    }
    const SynthJob J = jobs[j];
    const uint32_t local = blockIdx.x - J.wg_start;
    if (local == 0 && J.tail_zero)
        for (uint32_t k = threadIdx.x; k < J.n_tail; k += 256) J.tail_zero[k] = 0ull;
    synth_depth_tile(J.seed, J.w, J.h, J.occluders, J.p11, J.znear, J.depth, local % J.tiles_x, local / J.tiles_x,
                     s_rect, s_depth, &s_n);
}

}  // namespace

hipError_t launch_pyramid(const PyrBuildJob& job, uint32_t* err, hipStream_t st)
{
    if (job.w == 2u * job.base_w && job.h == 2u * job.base_h)
        hipLaunchKernelGGL(pyramid_kernel<kPyrBlock>, dim3(job.tiles_x * job.tiles_y), dim3(kPyrBlock), 0, st, job, err);
    else
        hipLaunchKernelGGL(pyramid_kernel<kPyrBlockWide>, dim3(job.tiles_x * job.tiles_y), dim3(kPyrBlockWide), 0, st,
                           job, err);
    return hipGetLastError();
}

hipError_t launch_pyramid_batch(const PyrBuildJob* jobs_dev, uint32_t n_jobs, uint32_t max_tiles, uint32_t* err,
                                hipStream_t st)
{
    if (n_jobs == 0 || max_tiles == 0) return hipSuccess;
    constexpr uint32_t kMaxGridY = 65535;  // jobs per launch (the grid's y extent); more go in further launches
    for (uint32_t j0 = 0; j0 < n_jobs; j0 += kMaxGridY) {
        const uint32_t n = n_jobs - j0 < kMaxGridY ? n_jobs - j0 : kMaxGridY;
        hipLaunchKernelGGL(pyramid_batch_kernel, dim3(max_tiles, n), dim3(kPyrBlock), 0, st, jobs_dev + j0, err);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_pyramid_level(const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw, uint32_t dh,
                                hipStream_t st)
{
    hipLaunchKernelGGL(pyramid_level_kernel, dim3((dw + 15) / 16, (dh + 15) / 16), dim3(256), 0, st, src, sw, sh, dst,
                       dw, dh);
    return hipGetLastError();
}

hipError_t launch_synth_depth_batch(const SynthJob* jobs_dev, const uint32_t* starts_dev, uint32_t n_jobs,
                                    uint32_t total_wgs, hipStream_t st)
{
    if (n_jobs == 0 || total_wgs == 0) return hipSuccess;
    hipLaunchKernelGGL(synth_depth_batch_kernel, dim3(total_wgs), dim3(256), 0, st, jobs_dev, starts_dev, n_jobs);
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
