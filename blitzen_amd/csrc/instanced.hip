// instanced.hip — LOD-instanced draw output for gfx950 (the D3D12 renderer's instanced cull).
//
// Replaces drawInstCountReset.cs.hlsl, drawInstCull.cs.hlsl and drawInstCmd.cs.hlsl (host dx12Draw.cpp:340-413):
// every visible object (frustum + LOD, decided by the decide kernel as for BZ_PASS_PLAIN) joins the instance list
// of its LOD, and each LOD with instances gets one instanced DrawIndexed command. The reference appends with
// InterlockedAdd (nondeterministic order, and a full list overflows into the next LOD's); here each list holds its
// objects in ascending objectId, at most per_lod of them, and the counter keeps the exact count:
//   hist:    per chunk of decided tiles, the emitting objects' count per LOD: counted by the decide kernel itself in
//            LDS when the LODs fit kInstLdsBins (a chunk is then one workgroup's contiguous run of tiles, and the
//            decided entries are read once, by the scatter); otherwise one workgroup per chunk x LOD slice builds an
//            LDS histogram from the decided entries (loaded as 64-entry groups, kBatch groups per round trip);
//   scan:    one workgroup per 64 LODs x segment of 64 chunks: exclusive prefix of the histogram over the chunks (the
//            chunk's first slot in the LOD's list; the segments' totals pass between workgroups by look-back) and the
//            counter {lod * per_lod, total};
//   cmds:    one extra workgroup of the scatter launch: the LODs with instances in ascending order -> 32-byte commands
//            and their number;
//   scatter: one workgroup per chunk x LOD slice: per-wave LOD counts over quarters of the chunk's entries, a hand-over
//            of each LOD's running slot through the quarters, then ranks among equal LODs with one ballot per bit of
//            the LOD index; each object written to lod * per_lod + slot.
#include "bz_internal.h"

namespace bz {

namespace {

constexpr uint32_t kSlice = 8192;  // LOD bins per histogram workgroup (LDS counters)
constexpr uint32_t kBatch = 8;     // 64-entry groups loaded per round trip

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

// Up to 64 consecutive decided tiles seen as one sequence of 64-entry groups (a tile's last group may be partial):
// lane k holds tile k's entry count and the inclusive prefix of the group counts.
struct TileWindow {
    uint32_t n, gincl, gex, groups;  // per lane (tile t0 + lane); groups = the window's total (wave-uniform)
};

__device__ __forceinline__ TileWindow load_window(const uint32_t* counts, uint32_t t0, uint32_t nt, uint32_t lane)
{
    TileWindow w;
    w.n = lane < nt ? counts[t0 + lane] : 0u;
    const uint32_t g = (w.n + 63u) >> 6;
    w.gincl = wave_incl_scan(g, lane);
    w.gex = w.gincl - g;
    w.groups = __shfl(w.gincl, 63, 64);
    return w;
}

// Entry `lane` of group g (< w.groups, wave-uniform) of the window: its tile is the first whose inclusive group
// prefix exceeds g. Returns false for the unused lanes of a partial group.
__device__ __forceinline__ bool load_group_entry(const TileWindow& w, uint32_t g, const uint2* staging, uint32_t t0,
                                                 uint32_t tile_size, uint32_t lane, uint2& out)
{
    const uint32_t j = (uint32_t)__popcll(__ballot(w.gincl <= g));
    const uint32_t nj = __shfl(w.n, j, 64), exj = __shfl(w.gex, j, 64);
    const uint32_t e = (g - exj) * 64u + lane;
    const bool valid = e < nj;
    out = valid ? staging[(size_t)(t0 + j) * tile_size + e] : make_uint2(0u, 0u);
    return valid;
}

// One workgroup (4 waves) per chunk of decided tiles x LOD slice: an LDS histogram of the chunk's LODs. Each wave
// takes batches of kBatch groups, all loaded before the first count.
__global__ __launch_bounds__(256) void inst_hist_kernel(InstancedJob J, const uint32_t* counts, const uint2* staging,
                                                        uint32_t* hist)
{
    __shared__ uint32_t h[kSlice];
    const uint32_t c = blockIdx.x, l0 = blockIdx.y * kSlice, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t l1 = l0 + kSlice < J.n_lods ? l0 + kSlice : J.n_lods;
    for (uint32_t i = tid; i < l1 - l0; i += 256) h[i] = 0u;
    __syncthreads();
    uint32_t c0, c1;
    chunk_range(c, J.tiles_per_chunk, J.n_long, J.n_tiles, c0, c1);
    for (uint32_t t0 = c0; t0 < c1; t0 += 64) {
        const uint32_t nt = c1 - t0 < 64u ? c1 - t0 : 64u;
        const TileWindow w = load_window(counts, t0, nt, lane);
        for (uint32_t gb = wave * kBatch; gb < w.groups; gb += 4u * kBatch) {
            uint2 ent[kBatch];
            bool val[kBatch];
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b) {
                val[b] = false;
                if (gb + b < w.groups) val[b] = load_group_entry(w, gb + b, staging, t0, J.tile_size, lane, ent[b]);
            }
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b)
                if (val[b] && ent[b].y >= l0 && ent[b].y < l1) atomicAdd(&h[ent[b].y - l0], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < l1 - l0; i += 256) hist[(size_t)c * J.n_lods + l0 + i] = h[i];
}

// Scan of the chunk x LOD histogram over the chunks, per LOD, by workgroups of (64 LODs: one per lane) x (segment of
// kSegChunks chunks: kSegChunks / 4 consecutive chunks per wave), so that a scan of ~1 M counters has hundreds of
// workgroups. A workgroup publishes its segment's total per LOD in one 64-bit word with a presence flag (value and
// flag in one atomic store: no fence), and sums the totals of the LOD group's earlier segments
// (look-back over at most a few dozen words, all loaded at once; a word of an earlier segment whose workgroup has not
// stored it yet is polled, with a bounded wait). Workgroups are dispatched in order of their linear index, segment
// major, so a workgroup waits only on workgroups dispatched before it. Each chunk's count is replaced by its first
// slot in the LOD's list; the last segment writes the counters {lod * per_lod, total}.
constexpr uint32_t kSegChunks = kInstSegChunks, kSegPerWave = kSegChunks / 4;

__global__ __launch_bounds__(256) void inst_scan_kernel(InstancedJob J, uint32_t* hist, unsigned long long* seg)
{
    __shared__ uint32_t wsum[4][64];
    __shared__ uint32_t before_s[64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t l = blockIdx.x * 64u + lane, s = blockIdx.y, n_seg = gridDim.y;
    const bool ok = l < J.n_lods;
    const uint32_t c0 = s * kSegChunks + wave * kSegPerWave;
    uint32_t v[kSegPerWave];
#pragma unroll
    for (uint32_t j = 0; j < kSegPerWave; ++j)
        v[j] = (ok && c0 + j < J.n_chunks) ? hist[(size_t)(c0 + j) * J.n_lods + l] : 0u;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kSegPerWave; ++j) sum += v[j];
    wsum[wave][lane] = sum;
    __syncthreads();
    uint32_t run = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
        const uint32_t x = wsum[w][lane];
        run += w < wave ? x : 0u;
        total += x;
    }
    constexpr unsigned long long tag = 1ull << 63;  // present; the scatter kernel clears the words for the next call
    if (wave == 0) {
        if (ok && s + 1u < n_seg)
            __hip_atomic_store(&seg[(size_t)s * J.n_lods + l], tag | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t before = 0;
        for (uint32_t k0 = 0; k0 < s; k0 += 16) {
            unsigned long long w[16];
#pragma unroll
            for (uint32_t k = 0; k < 16; ++k)
                w[k] = (ok && k0 + k < s) ? __hip_atomic_load(&seg[(size_t)(k0 + k) * J.n_lods + l], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
                                          : tag;
#pragma unroll
            for (uint32_t k = 0; k < 16; ++k) {
                uint32_t spins = 0;
                while (!(w[k] & tag)) {
                    if (++spins > J.spin_limit) {
                        atomicExch(J.err, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    w[k] = __hip_atomic_load(&seg[(size_t)(k0 + k) * J.n_lods + l], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
                }
                before += (uint32_t)w[k];
            }
        }
        before_s[lane] = before;
        if (ok && s + 1u == n_seg) {
            J.counters[2u * l] = l * J.per_lod;
            J.counters[2u * l + 1u] = before + total;
        }
    }
    __syncthreads();
    run += before_s[lane];
#pragma unroll
    for (uint32_t j = 0; j < kSegPerWave; ++j) {
        if (ok && c0 + j < J.n_chunks) hist[(size_t)(c0 + j) * J.n_lods + l] = run;
        run += v[j];
    }
}

// Commands for the LODs with instances, ascending (drawInstCmd.cs.hlsl's fields), and their number, by one workgroup
// of NT threads: thread t takes the contiguous LODs [t * R, (t + 1) * R), counts those with instances (their counters
// loaded 8 at a time), a block scan gives its first command slot, and it writes its commands in order.
// wsum: NT / 64 LDS words.
template <uint32_t NT>
__device__ __forceinline__ void build_commands(const InstancedJob& J, uint32_t* wsum)
{
    constexpr uint32_t NW = NT / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t R = (J.n_lods + NT - 1u) / NT;
    const uint32_t lb = tid * R < J.n_lods ? tid * R : J.n_lods;
    const uint32_t le = lb + R < J.n_lods ? lb + R : J.n_lods;
    uint32_t mine = 0;
    for (uint32_t l0 = lb; l0 < le; l0 += 8) {
        uint32_t c[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) c[j] = l0 + j < le ? J.counters[2u * (l0 + j) + 1u] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) mine += c[j] != 0u ? 1u : 0u;
    }
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t pos = incl - mine, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
        pos += w < wave ? wsum[w] : 0u;
        total += wsum[w];
    }
    if (tid == 0) *J.count = total;
    for (uint32_t l0 = lb; l0 < le; l0 += 8) {
        uint32_t c[8];
        uint2 d[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) c[j] = l0 + j < le ? J.counters[2u * (l0 + j) + 1u] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) d[j] = c[j] != 0u ? J.lod_draw[l0 + j] : make_uint2(0u, 0u);
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            if (c[j] == 0u) continue;
            if (pos < J.capacity) {
                const uint32_t l = l0 + j;
                uint4* cmd = reinterpret_cast<uint4*>(J.cmds + (size_t)pos * 8u);
                cmd[0] = make_uint4(l * J.per_lod, d[j].x, c[j] < J.per_lod ? c[j] : J.per_lod, d[j].y);  // objId = offset
                cmd[1] = make_uint4(0u, 0u, 0u, 0u);  // vertexOffset, firstInstance, padding
            }
            ++pos;
        }
    }
}

// The commands alone (a range without objects: no scatter launch follows).
__global__ __launch_bounds__(1024) void inst_cmd_kernel(InstancedJob J)
{
    __shared__ uint32_t wsum[16];
    build_commands<1024>(J, wsum);
}

// One workgroup (4 waves) per chunk x LOD slice of kScatterSlice LODs. The chunk's entries form one sequence of
// 64-entry groups (ascending objectId) per window of up to 64 tiles; wave w takes the w-th quarter of the groups.
//   1. each wave counts its quarter's LODs into its own LDS histogram;
//   2. per LOD, the running slot (starting at the chunk's first slot in the LOD's list, from the scanned histogram)
//      is handed through the four quarters in order: each wave's histogram becomes its first slot per LOD;
//   3. each wave ranks its entries among equal LODs in its quarter, in order: the lanes holding the same LOD are found
//      with one ballot per bit of the LOD's index in the slice, and the lowest such lane advances the LOD's slot.
constexpr uint32_t kScatterSlice = 2048;
constexpr uint32_t kScatterWaves = 4;

__global__ __launch_bounds__(kScatterWaves * 64) void inst_scatter_kernel(InstancedJob J, const uint32_t* counts,
                                                                          const uint2* staging,
                                                                          const uint32_t* hist,
                                                                          unsigned long long* seg, uint64_t n_seg_words)
{
    {  // the scan's published words, cleared for the next call (the scan has finished: this is the next launch)
        const uint64_t nthreads = (uint64_t)gridDim.x * gridDim.y * blockDim.x;
        for (uint64_t i = ((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x; i < n_seg_words;
             i += nthreads)
            __hip_atomic_store(&seg[i], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __shared__ uint32_t slot[kScatterWaves][kScatterSlice];  // per wave: LOD -> count (1), then next slot (2, 3)
    __shared__ uint32_t running[kScatterSlice];              // per LOD: the next slot after the windows so far
    __shared__ uint32_t cmd_wsum[kScatterWaves];
    if (blockIdx.x == J.n_chunks) {  // one extra workgroup writes the commands (final counters since the scan)
        if (blockIdx.y == 0) build_commands<kScatterWaves * 64>(J, cmd_wsum);
        return;
    }
    const uint32_t c = blockIdx.x, l0 = blockIdx.y * kScatterSlice, tid = threadIdx.x, lane = tid & 63u,
                   wave = tid >> 6;
    const uint32_t l1 = l0 + kScatterSlice < J.n_lods ? l0 + kScatterSlice : J.n_lods;
    const uint32_t nl = l1 - l0;
    const uint32_t nbits = nl > 1u ? 32u - (uint32_t)__clz(nl - 1u) : 0u;  // bits of the slice's largest key
    const unsigned long long below = (1ull << lane) - 1ull;
    {
        const uint32_t* row = hist + (size_t)c * J.n_lods + l0;
        uint32_t v[kScatterSlice / (kScatterWaves * 64)];
#pragma unroll
        for (uint32_t k = 0; k < kScatterSlice / (kScatterWaves * 64); ++k) {
            const uint32_t i = k * kScatterWaves * 64u + tid;
            v[k] = i < nl ? row[i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kScatterSlice / (kScatterWaves * 64); ++k) {
            const uint32_t i = k * kScatterWaves * 64u + tid;
            if (i < nl) running[i] = v[k];
        }
    }
    uint32_t c0, c1;
    chunk_range(c, J.tiles_per_chunk, J.n_long, J.n_tiles, c0, c1);
    for (uint32_t t0 = c0; t0 < c1; t0 += 64) {
        const uint32_t nt = c1 - t0 < 64u ? c1 - t0 : 64u;
        const TileWindow w = load_window(counts, t0, nt, lane);
        const uint32_t q = (w.groups + kScatterWaves - 1u) / kScatterWaves;  // groups per wave
        const uint32_t g0 = wave * q < w.groups ? wave * q : w.groups;
        const uint32_t g1 = g0 + q < w.groups ? g0 + q : w.groups;
        for (uint32_t i = lane; i < nl; i += 64) slot[wave][i] = 0u;
        __syncthreads();  // (also orders `running` from the previous window or the seeding)
        // 1. this wave's LOD counts
        uint2 ent[kBatch];
        bool val[kBatch];
        for (uint32_t gb = g0; gb < g1; gb += kBatch) {
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b) {
                val[b] = false;
                if (gb + b < g1) val[b] = load_group_entry(w, gb + b, staging, t0, J.tile_size, lane, ent[b]);
            }
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b)
                if (val[b] && ent[b].y >= l0 && ent[b].y < l1) atomicAdd(&slot[wave][ent[b].y - l0], 1u);
        }
        __syncthreads();
        // 2. hand the running slot through the waves' quarters, per LOD
        for (uint32_t i = tid; i < nl; i += kScatterWaves * 64u) {
            uint32_t run = running[i];
#pragma unroll
            for (uint32_t v = 0; v < kScatterWaves; ++v) {
                const uint32_t n = slot[v][i];
                slot[v][i] = run;
                run += n;
            }
            running[i] = run;
        }
        __syncthreads();
        // 3. ranks in order within this wave's quarter (the entries of the last batch are still in registers when the
        // quarter fits one batch)
        const bool reuse = g1 - g0 <= kBatch;
        for (uint32_t gb = g0; gb < g1; gb += kBatch) {
            if (!reuse) {
#pragma unroll
                for (uint32_t b = 0; b < kBatch; ++b) {
                    val[b] = false;
                    if (gb + b < g1) val[b] = load_group_entry(w, gb + b, staging, t0, J.tile_size, lane, ent[b]);
                }
            }
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b) {
                if (gb + b >= g1) break;
                const bool in = val[b] && ent[b].y >= l0 && ent[b].y < l1;
                const uint32_t key = in ? ent[b].y - l0 : 0u;
                unsigned long long eq = __ballot(in);  // lanes in the slice with this lane's key
                for (uint32_t bit = 0; bit < nbits; ++bit) {
                    const bool set = (key >> bit) & 1u;
                    const unsigned long long m = __ballot(set);
                    eq &= set ? m : ~m;
                }
                const uint32_t rank = (uint32_t)__popcll(eq & below);
                const uint32_t pos = in ? slot[wave][key] + rank : 0u;  // every lane reads before any leader writes
                __builtin_amdgcn_wave_barrier();
                if (in && rank == 0u) slot[wave][key] = pos + (uint32_t)__popcll(eq);
                __builtin_amdgcn_wave_barrier();
                if (in && pos < J.per_lod) J.instance_ids[(size_t)ent[b].y * J.per_lod + pos] = ent[b].x;
            }
        }
        __syncthreads();  // slot[] is cleared by the next window
    }
}

}  // namespace

hipError_t launch_instanced(const InstancedJob& J, const CompactScratch& s, uint32_t* hist, unsigned long long* seg,
                            bool with_hist, hipStream_t st)
{
    if (J.n_chunks && !with_hist) {
        const uint32_t slices = (J.n_lods + kSlice - 1) / kSlice;
        hipLaunchKernelGGL(inst_hist_kernel, dim3(J.n_chunks, slices), dim3(256), 0, st, J,
                           (const uint32_t*)s.counts, (const uint2*)s.staging, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(inst_scan_kernel, dim3((J.n_lods + 63) / 64, inst_segments(J.n_chunks)), dim3(256), 0, st, J,
                       hist, seg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (!J.n_chunks) {
        hipLaunchKernelGGL(inst_cmd_kernel, dim3(1), dim3(1024), 0, st, J);
        return hipGetLastError();
    }
    const uint32_t scatter_slices = (J.n_lods + kScatterSlice - 1) / kScatterSlice;
    hipLaunchKernelGGL(inst_scatter_kernel, dim3(J.n_chunks + 1u, scatter_slices), dim3(kScatterWaves * 64), 0, st, J,
                       (const uint32_t*)s.counts, (const uint2*)s.staging, (const uint32_t*)hist, seg,
                       (uint64_t)(inst_segments(J.n_chunks) - 1u) * J.n_lods);
    return hipGetLastError();
}

}  // namespace bz
