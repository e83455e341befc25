// cluster.hip — gfx950 kernels of the two-level cluster cull's second level.
//
// Replaces the expansion of PreClusterDrawCull.comp.glsl:13-47 (atomicAdd(clusterCount, lod.clusterCount), one
// ClusterDispatchData per cluster) and InitialClusterCull.comp.glsl:12-55 (one draw per cluster). The reference
// reads the expanded count back to the host between the passes (vulkanDraw.cpp:874-905); here everything stays on
// the device and the order is deterministic:
//   object pass (cull.hip, decide_cluster_kernel): per object tile, the emitting objects (objectId, LOD) in object
//       order and the tile's cluster total;
//   expand:  one workgroup per object tile. Its first record is the sum of the earlier tiles' totals (added by the
//       workgroup itself for up to kScanTiles tiles, else read from a scan of the totals launched before); it writes the
//       tile's records in (objectId, clusterId) order, tests every record's cluster (frustum, backface cone, Hi-Z)
//       and compacts the survivors, in record order, into the tile's own record range of a survivor buffer;
//   emit:    one wave per tile adds the earlier tiles' survivor counts and writes the DrawIndexedIndirect commands
//       {objectId, triangleCount*3, 1, dataOffset, 0, 0} and the draw count.
// Cluster tests are this library's definition (the reference leaves them commented out,
// InitialClusterCull.comp.glsl:24-42): the cluster sphere goes through the object transform and the view like an
// object sphere; the cone is culled when dot(c, a) >= cutoff * |c| + r with c the view-space centre, a the cone
// axis (int8 / 127) rotated by the object quaternion and the view's 3x3, cutoff = int8 / 127 (the camera sits at
// the view-space origin; meshoptimizer's sphere-corrected cone test).
#include "bz_internal.h"

namespace bz {

namespace {

constexpr uint32_t kXBlock = 256;
constexpr uint32_t kXWaves = kXBlock / 64;
constexpr uint32_t kXMaxTile = 2048;  // largest object tile (block x items) of the object pass

struct ExpandSmem {
    uint32_t incl[kXMaxTile];  // inclusive prefix of the tile's emitting objects' cluster counts
    uint32_t obj[kXMaxTile];
    uint32_t lod[kXMaxTile];
    uint32_t coff[kXMaxTile];  // first cluster of the object's LOD
    uint32_t red[kXWaves];
    uint32_t lvl[kMaxMips];
};

__device__ __forceinline__ uint32_t incl_scan64(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of a[u] for u in [0, end), by the whole workgroup (all loads of a thread issued before the first add).
__device__ __forceinline__ uint32_t block_prefix(const uint32_t* a, uint32_t end, uint32_t* red)
{
    const uint32_t tid = threadIdx.x;
    uint32_t part = 0;
    for (uint32_t u0 = 0; u0 < end; u0 += 16u * kXBlock) {
        uint32_t v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t u = u0 + (uint32_t)j * kXBlock + tid;
            v[j] = u < end ? a[u] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) part += v[j];
    }
    part = wave_sum(part);
    if ((tid & 63u) == 0) red[tid >> 6] = part;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (uint32_t w = 0; w < kXWaves; ++w) s += red[w];
    __syncthreads();  // red is reused by the caller
    return s;
}

__device__ __forceinline__ float int8_unit(uint32_t packed, uint32_t shift)
{
    return (float)(int8_t)(uint8_t)(packed >> shift) / 127.0f;
}

// The cluster tests of one record (same arithmetic as the object cull, cull_math.h).
__device__ __forceinline__ bool cluster_visible(const ClusterJob& J, uint32_t obj, uint32_t cid, const uint32_t* lvl)
{
    const SceneDev& S = J.scene;
    const uint32_t x = S.obj_xf ? S.obj_xf[obj] : obj;
    const float4 ps = S.xf_ps[x], q = S.xf_q[x], sp = J.cl_sphere[cid];
    float cx, cy, cz, r;
    sphere_to_view(J.view, sp.x, sp.y, sp.z, sp.w, ps.x, ps.y, ps.z, ps.w, q.x, q.y, q.z, q.w, cx, cy, cz, r);
    bool vis = true;
    if (J.tests & BZ_CLUSTER_FRUSTUM) vis = frustum_visible(J.view, cx, cy, cz, r);
    if (vis && (J.tests & BZ_CLUSTER_CONE)) {
        const uint32_t cone = J.cl_cone[cid];
        float wx, wy, wz;
        rotate_quat(int8_unit(cone, 0), int8_unit(cone, 8), int8_unit(cone, 16), q.x, q.y, q.z, q.w, wx, wy, wz);
        // the view's 3x3 (no translation): ((view[k]*x + view[4+k]*y) + view[8+k]*z), view[c*4+k] = m[4k+c]
        const float* m = J.view.m;
        const float ax = (m[0] * wx + m[1] * wy) + m[2] * wz;
        const float ay = (m[4] * wx + m[5] * wy) + m[6] * wz;
        const float az = (m[8] * wx + m[9] * wy) + m[10] * wz;
        const float cutoff = int8_unit(cone, 24);
        const float d = (cx * ax + cy * ay) + cz * az;
        const float len = sqrtf((cx * cx + cy * cy) + cz * cz);
        if (d >= cutoff * len + r) vis = false;
    }
    if (vis && (J.tests & BZ_CLUSTER_HIZ) && J.hiz) {
        float a0, a1, a2, a3;
        if (project_sphere(cx, cy, cz, r, J.view.znear, J.view.p00, J.view.p11, a0, a1, a2, a3))
            vis = occlusion_visible(J.view, J.pyr, lvl, a0, a1, a2, a3, cz, r);
    }
    return vis;
}

__global__ __launch_bounds__(kXBlock) void cluster_expand_kernel(ClusterJob J, const uint32_t* counts,
                                                                 const uint2* staging, const uint32_t* ccounts,
                                                                 uint32_t* rbase, uint32_t* scount, uint2* survivors)
{
    __shared__ ExpandSmem sm;
    const uint32_t t = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t ne = counts[t];
    // the tile's first kXBlock staging entries, loaded with its count (before it is known; every tile owns
    // tile_size >= 2 * kXBlock slots), so that they arrive in the same round trip as the counts and the prefix
    const uint2 first_en = staging[(size_t)t * J.tile_size + tid];
    if (tid < (uint32_t)kMaxMips) {  // pyramid level offsets from the extent
        uint32_t off = 0;
        for (uint32_t L = 0; L < tid; ++L) off += level_dim(J.pyr.base_w, L) * level_dim(J.pyr.base_h, L);
        sm.lvl[tid] = off;
    }
    const uint32_t base = J.scanned ? rbase[t] : block_prefix(ccounts, t, sm.red);

    // the tile's emitting objects -> LOD cluster ranges, inclusive scan of the cluster counts
    uint32_t carry = 0;
    for (uint32_t e0 = 0; e0 < ne; e0 += kXBlock) {
        const uint32_t e = e0 + tid;
        uint32_t cnt = 0;
        if (e < ne) {
            const uint2 en = e0 == 0 ? first_en : staging[(size_t)t * J.tile_size + e];
            const uint2 lc = J.scene.lod_cluster[en.y];
            sm.obj[e] = en.x;
            sm.lod[e] = en.y;
            sm.coff[e] = lc.x;
            cnt = lc.y;
        }
        const uint32_t incl = incl_scan64(cnt, lane);
        if (lane == 63) sm.red[wave] = incl;
        __syncthreads();
        uint32_t wbase = carry, chunk = 0;
#pragma unroll
        for (uint32_t w = 0; w < kXWaves; ++w) {
            wbase += w < wave ? sm.red[w] : 0u;
            chunk += sm.red[w];
        }
        if (e < ne) sm.incl[e] = wbase + incl;
        carry += chunk;
        __syncthreads();  // red is rewritten by the next chunk
    }
    const uint32_t T = carry;  // == ccounts[t]
    if (tid == 0) {
        if (!J.scanned) rbase[t] = base;
        if (t == J.n_tiles - 1u) *J.dispatch_count = base + T;
    }

    // records in (objectId, clusterId) order; survivors compacted in the same order
    uint32_t surv = 0;
    for (uint32_t r0 = 0; r0 < T; r0 += kXBlock) {
        const uint32_t r = r0 + tid;
        const uint64_t gr = (uint64_t)base + r;
        const bool act = r < T && gr < J.dispatch_capacity;
        bool keep = false;
        uint32_t obj = 0, cid = 0;
        if (act) {
            uint32_t lo = 0, hi = ne - 1u;  // first entry with incl > r (exists: r < T = incl[ne-1])
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sm.incl[mid] > r) hi = mid;
                else lo = mid + 1u;
            }
            const uint32_t first = lo ? sm.incl[lo - 1u] : 0u;
            obj = sm.obj[lo];
            cid = sm.coff[lo] + (r - first);
            uint32_t* d = J.dispatch + gr * 3u;
            d[0] = obj;
            d[1] = sm.lod[lo];
            d[2] = cid;
            keep = cluster_visible(J, obj, cid, sm.lvl);
        }
        const unsigned long long m = __ballot(keep);
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) sm.red[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = surv, chunk = 0;
#pragma unroll
        for (uint32_t w = 0; w < kXWaves; ++w) {
            off += w < wave ? sm.red[w] : 0u;
            chunk += sm.red[w];
        }
        if (keep) survivors[(size_t)base + off + rank] = make_uint2(obj, cid);  // <= gr < dispatch_capacity
        surv += chunk;
        __syncthreads();
    }
    if (tid == 0) scount[t] = surv;
}

constexpr uint32_t kCEmitTiles = kXBlock / 64;
constexpr uint32_t kCEmitUnroll = 4;

__global__ __launch_bounds__(kXBlock) void cluster_emit_kernel(ClusterJob J, const uint32_t* rbase,
                                                               const uint32_t* scount, const uint32_t* sprefix,
                                                               const uint2* survivors)
{
    __shared__ uint32_t red[kXWaves];
    __shared__ uint32_t cnt[kCEmitTiles + 1];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t t0 = blockIdx.x * kCEmitTiles;
    const uint32_t t1 = t0 + kCEmitTiles < J.n_tiles ? t0 + kCEmitTiles : J.n_tiles;
    const uint32_t nt = t1 - t0;
    const uint32_t own = tid < nt ? scount[t0 + tid] : 0u;
    const uint32_t prefix = J.scanned ? sprefix[t0] : block_prefix(scount, t0, red);
    if (wave == 0) {
        uint32_t v = own;
#pragma unroll
        for (int off = 1; off < (int)kCEmitTiles; off <<= 1) {
            const uint32_t x = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off) v += x;
        }
        if (lane < kCEmitTiles) cnt[lane + 1] = v;
        if (lane == 0) cnt[0] = 0;
    }
    __syncthreads();
    if (tid == 0 && t1 == J.n_tiles) *J.count = prefix + cnt[nt];
    if (wave >= nt) return;
    const uint32_t c = cnt[wave + 1] - cnt[wave];
    const uint32_t out0 = prefix + cnt[wave];
    const uint2* src = survivors + rbase[t0 + wave];
    for (uint32_t e0 = 0; e0 < c; e0 += 64u * kCEmitUnroll) {
        uint2 ent[kCEmitUnroll], d[kCEmitUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kCEmitUnroll; ++u) {
            const uint32_t e = e0 + u * 64u + lane;
            ent[u] = src[e < c ? e : 0u];  // objectId, clusterId (entry 0 exists whenever c > 0)
        }
#pragma unroll
        for (uint32_t u = 0; u < kCEmitUnroll; ++u) d[u] = J.cl_draw[ent[u].y];
#pragma unroll
        for (uint32_t u = 0; u < kCEmitUnroll; ++u) {
            const uint32_t e = e0 + u * 64u + lane;
            const uint32_t pos = out0 + e;
            if (e < c && pos < J.capacity) {
                uint2* cmd = reinterpret_cast<uint2*>(J.cmds + (size_t)pos * 6u);
                cmd[0] = make_uint2(ent[u].x, d[u].x);  // objectId, indexCount
                cmd[1] = make_uint2(1u, d[u].y);         // instanceCount, firstIndex
                cmd[2] = make_uint2(0u, 0u);             // vertexOffset, firstInstance
            }
        }
    }
}

}  // namespace

hipError_t launch_cluster_pass(const ClusterJob& job_in, const CompactScratch& cs, const ClusterScratch& xs,
                               hipStream_t st)
{
    if (job_in.n_tiles == 0 || job_in.tile_size > kXMaxTile) return hipErrorInvalidValue;
    ClusterJob job = job_in;
    job.scanned = job.n_tiles > kScanTiles ? 1u : 0u;
    hipError_t e = hipSuccess;
    if (job.scanned) e = launch_scan_counts(xs.ccounts, job.n_tiles, xs.rbase, st);  // record bases
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cluster_expand_kernel, dim3(job.n_tiles), dim3(kXBlock), 0, st, job,
                       (const uint32_t*)cs.counts, (const uint2*)cs.staging, (const uint32_t*)xs.ccounts, xs.rbase,
                       xs.scount, xs.survivors);
    e = hipGetLastError();
    if (e == hipSuccess && job.scanned) e = launch_scan_counts(xs.scount, job.n_tiles, xs.sprefix, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cluster_emit_kernel, dim3((job.n_tiles + kCEmitTiles - 1) / kCEmitTiles), dim3(kXBlock), 0, st,
                       job, (const uint32_t*)xs.rbase, (const uint32_t*)xs.scount, (const uint32_t*)xs.sprefix,
                       (const uint2*)xs.survivors);
    return hipGetLastError();
}

}  // namespace bz
