// instanced.hip — LOD-instanced draw output for gfx950 (the D3D12 renderer's instanced cull).
//
// Replaces drawInstCountReset.cs.hlsl, drawInstCull.cs.hlsl and drawInstCmd.cs.hlsl (host dx12Draw.cpp:340-413):
// every visible object (frustum + LOD, decided by the decide kernel as for BZ_PASS_PLAIN) joins the instance list
// of its LOD, and each LOD with instances gets one instanced DrawIndexed command. The reference appends with
// InterlockedAdd (nondeterministic order, and a full list overflows into the next LOD's); here each list holds its
// objects in ascending objectId, at most per_lod of them, and the counter keeps the exact count:
//   hist:    one workgroup per chunk of decided tiles x LOD slice: an LDS histogram of the chunk's LODs;
//   scan:    one thread per LOD: exclusive prefix of the histogram over chunks (the chunk's first slot in the LOD's
//            list) and the counter {lod * per_lod, total};
//   cmds:    one workgroup: the LODs with instances in ascending order -> 32-byte commands and their number;
//   scatter: one wave per chunk x LOD slice: the chunk's objects in order, ranked among equal LODs by peeling
//            (one ballot per distinct LOD in a group of 64), written to lod * per_lod + slot.
#include "bz_internal.h"

namespace bz {

namespace {

constexpr uint32_t kSlice = 8192;  // LOD bins per histogram / scatter workgroup (LDS counters)

__global__ __launch_bounds__(256) void inst_hist_kernel(InstancedJob J, const uint32_t* counts, const uint2* staging,
                                                        uint32_t* hist)
{
    __shared__ uint32_t h[kSlice];
    const uint32_t c = blockIdx.x, l0 = blockIdx.y * kSlice, tid = threadIdx.x;
    const uint32_t l1 = l0 + kSlice < J.n_lods ? l0 + kSlice : J.n_lods;
    for (uint32_t i = tid; i < l1 - l0; i += 256) h[i] = 0u;
    __syncthreads();
    const uint32_t t0 = c * J.tiles_per_chunk;
    const uint32_t t1 = t0 + J.tiles_per_chunk < J.n_tiles ? t0 + J.tiles_per_chunk : J.n_tiles;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t n = counts[t];
        const uint2* st = staging + (size_t)t * J.tile_size;
        for (uint32_t e = tid; e < n; e += 256) {
            const uint32_t lod = st[e].y;
            if (lod >= l0 && lod < l1) atomicAdd(&h[lod - l0], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < l1 - l0; i += 256) hist[(size_t)c * J.n_lods + l0 + i] = h[i];
}

// One thread per LOD: the histogram column becomes each chunk's first slot in the LOD's list (in place).
__global__ __launch_bounds__(256) void inst_scan_kernel(InstancedJob J, uint32_t* hist)
{
    const uint32_t l = blockIdx.x * 256 + threadIdx.x;
    if (l >= J.n_lods) return;
    uint32_t run = 0;
    uint32_t c = 0;
    for (; c + 8 <= J.n_chunks; c += 8) {  // eight independent loads in flight
        uint32_t v[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) v[j] = hist[(size_t)(c + j) * J.n_lods + l];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            hist[(size_t)(c + j) * J.n_lods + l] = run;
            run += v[j];
        }
    }
    for (; c < J.n_chunks; ++c) {
        const uint32_t v = hist[(size_t)c * J.n_lods + l];
        hist[(size_t)c * J.n_lods + l] = run;
        run += v;
    }
    J.counters[2u * l] = l * J.per_lod;
    J.counters[2u * l + 1u] = run;
}

// One workgroup: commands for the LODs with instances, ascending (drawInstCmd.cs.hlsl's fields).
__global__ __launch_bounds__(1024) void inst_cmd_kernel(InstancedJob J)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t s_base;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) s_base = 0;
    __syncthreads();
    for (uint32_t l0 = 0; l0 < J.n_lods; l0 += 1024) {
        const uint32_t l = l0 + tid;
        const uint32_t cnt = l < J.n_lods ? J.counters[2u * l + 1u] : 0u;
        const bool has = cnt != 0u;
        const unsigned long long m = __ballot(has);
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = s_base, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < 16; ++w) {
            before += w < wave ? wsum[w] : 0u;
            total += wsum[w];
        }
        const uint32_t pos = before + rank;
        if (has && pos < J.capacity) {
            const uint2 d = J.lod_draw[l];
            uint4* cmd = reinterpret_cast<uint4*>(J.cmds + (size_t)pos * 8u);
            cmd[0] = make_uint4(l * J.per_lod, d.x, cnt < J.per_lod ? cnt : J.per_lod, d.y);  // objId = offset
            cmd[1] = make_uint4(0u, 0u, 0u, 0u);  // vertexOffset, firstInstance, padding
        }
        __syncthreads();
        if (tid == 0) s_base += total;
        __syncthreads();
    }
    if (tid == 0) *J.count = s_base;
}

// One wave per chunk x LOD slice: the chunk's objects in ascending objectId, each ranked among the chunk's objects of
// the same LOD seen before it.
__global__ __launch_bounds__(64) void inst_scatter_kernel(InstancedJob J, const uint32_t* counts, const uint2* staging,
                                                          const uint32_t* base)
{
    __shared__ uint32_t cnt[kSlice];
    const uint32_t c = blockIdx.x, l0 = blockIdx.y * kSlice, lane = threadIdx.x;
    const uint32_t l1 = l0 + kSlice < J.n_lods ? l0 + kSlice : J.n_lods;
    for (uint32_t i = lane; i < l1 - l0; i += 64) cnt[i] = 0u;
    __syncthreads();
    const uint32_t t0 = c * J.tiles_per_chunk;
    const uint32_t t1 = t0 + J.tiles_per_chunk < J.n_tiles ? t0 + J.tiles_per_chunk : J.n_tiles;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t n = counts[t];
        const uint2* st = staging + (size_t)t * J.tile_size;
        for (uint32_t e0 = 0; e0 < n; e0 += 64) {
            const uint32_t e = e0 + lane;
            const uint2 en = e < n ? st[e] : make_uint2(0u, 0u);
            const bool in = e < n && en.y >= l0 && en.y < l1;
            unsigned long long active = __ballot(in);
            uint32_t rank = 0;
            while (active) {
                const uint32_t leader = (uint32_t)__ffsll((long long)active) - 1u;
                const uint32_t key = __shfl(en.y, leader, 64);
                const bool mine = in && en.y == key;
                const unsigned long long m = __ballot(mine);
                if (mine) rank = cnt[key - l0] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                __builtin_amdgcn_wave_barrier();
                if (lane == leader) cnt[key - l0] += (uint32_t)__popcll(m);
                __builtin_amdgcn_wave_barrier();
                active &= ~m;
            }
            if (in) {
                const uint32_t pos = base[(size_t)c * J.n_lods + en.y] + rank;
                if (pos < J.per_lod) J.instance_ids[(size_t)en.y * J.per_lod + pos] = en.x;
            }
        }
    }
}

}  // namespace

hipError_t launch_instanced(const InstancedJob& J, const CompactScratch& s, uint32_t* hist, hipStream_t st)
{
    const uint32_t slices = (J.n_lods + kSlice - 1) / kSlice;
    if (J.n_chunks) {
        hipLaunchKernelGGL(inst_hist_kernel, dim3(J.n_chunks, slices), dim3(256), 0, st, J,
                           (const uint32_t*)s.counts, (const uint2*)s.staging, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(inst_scan_kernel, dim3((J.n_lods + 255) / 256), dim3(256), 0, st, J, hist);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(inst_cmd_kernel, dim3(1), dim3(1024), 0, st, J);
    e = hipGetLastError();
    if (e != hipSuccess || !J.n_chunks) return e;
    hipLaunchKernelGGL(inst_scatter_kernel, dim3(J.n_chunks, slices), dim3(64), 0, st, J, (const uint32_t*)s.counts,
                       (const uint2*)s.staging, (const uint32_t*)hist);
    return hipGetLastError();
}

}  // namespace bz
