// bz_internal.h — device-side job descriptors and kernel launchers shared by the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bzcull.h"
#include "cull_math.h"

namespace bz {

// Resident scene in HBM, structure-of-arrays. Transforms are split into two 16-byte streams so every
// per-object load of a wavefront is one fully coalesced 1 KiB access.
struct SceneDev {
    const uint32_t* obj_surf;   // [n_objects] RenderObject.surfaceId (nullptr when stored as 16-bit)
    const uint16_t* obj_surf16; // [n_objects] RenderObject.surfaceId when every id < 65536 (else nullptr)
    const uint32_t* obj_xf;     // [n_objects] RenderObject.transformId; nullptr when transformId == objectId
    const float4* xf_ps;        // [n_transforms] pos.xyz, scale
    const float4* xf_q;         // [n_transforms] orientation
    const float4* surf_sphere;  // [n_surfaces] center.xyz, radius
    const uint2* surf_lod;      // [n_surfaces] lodOffset, lodCount
    const float* lod_error;     // [n_lods]
    const uint2* lod_draw;      // [n_lods] indexCount, firstIndex
    const uint2* lod_cluster;   // [n_lods] clusterOffset, clusterCount
    uint32_t n_objects, n_transforms, n_surfaces, n_lods;
};

enum : uint32_t { PASS_EARLY = 0, PASS_LATE = 1, PASS_PLAIN = 2, PASS_TEMPORAL = 3, PASS_ONPC = 4 };

// One view's cull: a kernel argument for the single-view kernel and one entry of the batched job table.
struct CullJob {
    SceneDev scene;
    ViewConsts view;
    PyramidDesc pyr;              // pyr.texels == nullptr: no Hi-Z test
    const uint32_t* vis_in;       // EARLY / LATE
    uint32_t* vis_out;            // LATE
    uint32_t* cmds;               // IndirectDrawData[capacity] as 6 dwords each
    uint32_t* count;
    uint32_t begin, n, capacity, pass;
    uint32_t hiz;                 // apply the Hi-Z test
    uint32_t tile_start;          // batched: first global tile of this job
    uint32_t n_tiles;             // tiles of this job
    uint32_t pad;
};

// Scratch of the two-kernel compaction: per-tile command counts and per-tile staging slots (objectId, LOD) of
// tile_size entries each, and room for the counts' exclusive prefix. Fully rewritten by every decide launch; nothing
// needs clearing between launches. `counts` and `prefix` share one allocation of counts_words(capacity_tiles).
struct CompactScratch {
    uint32_t* counts;            // [capacity_tiles]
    uint2* staging;              // [capacity_entries]
    uint32_t capacity_tiles;
    uint64_t capacity_entries;
    uint32_t* prefix;            // [capacity_tiles + 1] exclusive prefix of counts (written by the scan, when used)
};
constexpr uint64_t counts_words(uint32_t capacity_tiles) { return 2ull * capacity_tiles + 1u; }
inline void set_prefix(CompactScratch& s) { s.prefix = s.counts ? s.counts + s.capacity_tiles : nullptr; }

// An emit workgroup adds the counts of its job's earlier tiles itself when there are at most kScanTiles of them (one
// round of loads). A job with more tiles gets their exclusive prefix from one scan launch between the decide and the
// emit instead, so the emit's work per workgroup stays constant and the pass stays linear in the object count.
constexpr uint32_t kScanTiles = 4096;
// Exclusive prefix of counts[0, n) into prefix[0, n], prefix[n] = the total (one workgroup, linear in n).
hipError_t launch_scan_counts(const uint32_t* counts, uint32_t n, uint32_t* prefix, hipStream_t st);

struct PyrBuildJob {
    const float* depth;
    uint32_t w, h;
    float* texels;
    uint32_t base_w, base_h, mips;
    uint32_t tile_w, tile_h, tiles_x, tiles_y;  // single-pass tiling of level 0
    uint32_t tile_levels;                        // levels reduced inside a tile (>= 1: includes level 0)
    uint32_t wg_start;                           // batched: first workgroup of this job
    uint32_t spin_limit;                         // sleeps before a wait for a tail word gives up
    unsigned long long* tail;                    // per job: one flagged word per tile, zero between builds
    uint32_t level_offset[16];
};

constexpr int kMaxMips = 16;

// Launchers (cull.hip). Each returns hipSuccess or the launch error.
// The decide kernel is persistent: n_cu * per_cu workgroups (per_cu == 0: as many as are resident per CU),
// clamped to the tile count.
// Frame cache of a scene: the frustum survivors an EARLY pass found (view-space sphere, objectId, LOD index), per tile
// in ascending objectId order, for the LATE pass of the same view, transforms and object range to reuse. Nothing in
// it depends on the visibility buffer: the LATE pass reads the current visibility itself.
struct FrameCacheDev {
    float4* sph;       // [tiles * tile] cx, cy, cz, r
    uint2* meta;       // [tiles * tile] objectId, absolute LOD index
    uint32_t* counts;  // [tiles] survivors per tile
};

// cache (EARLY only, may be null): also test every object of the range against the frustum and record the
// survivors there.
hipError_t launch_cull(const CullJob& job, const CompactScratch& s, uint32_t block, uint32_t items, uint32_t n_cu,
                       uint32_t per_cu, hipStream_t st, const FrameCacheDev* cache = nullptr);
// Chunks of consecutive tiles: chunk c of a job holds tiles [t0, t1), the first n_long chunks tpc + 1 tiles and the
// others tpc (so that as many chunks as workgroups fit at once take the whole job evenly).
__host__ __device__ inline void chunk_range(uint32_t c, uint32_t tpc, uint32_t n_long, uint32_t n_tiles, uint32_t& t0,
                                            uint32_t& t1)
{
    t0 = c * tpc + (c < n_long ? c : n_long);
    t1 = t0 + tpc + (c < n_long ? 1u : 0u);
    if (t1 > n_tiles) t1 = n_tiles;
    if (t0 > t1) t0 = t1;
}
// LOD histogram per chunk of decided tiles, filled by the decide kernel of the LOD-instanced output: workgroup c
// decides the tiles of chunk c (chunk_range) in order, counts its emitting objects per LOD in LDS
// (n_lods <= kInstLdsBins) and writes the counts to hist[c * n_lods + LOD].
struct InstHistDev {
    uint32_t* hist;
    uint32_t tiles_per_chunk, n_long, n_lods;
};
// The decide kernel alone: per tile, the emitting objects' (objectId, LOD) in ascending objectId and their count
// (PLAIN decisions), and with `ih` their LOD histogram per chunk of tiles.
hipError_t launch_decide_only(const CullJob& job, const CompactScratch& s, uint32_t block, uint32_t items,
                              uint32_t n_cu, uint32_t per_cu, hipStream_t st, const InstHistDev* ih = nullptr);

// LOD-instanced output from a decided job (tiles of tile_size entries in s): per LOD, the ascending object ids
// (at most per_lod) at lod * per_lod, the counters {lod * per_lod, exact count}, and one 32-byte command per LOD
// with instances, in ascending LOD order, plus their number.
struct InstancedJob {
    const uint2* lod_draw;        // [n_lods] indexCount, firstIndex
    uint32_t* instance_ids;       // [n_lods * per_lod]
    uint32_t* counters;           // [n_lods] x 2
    uint32_t* cmds;               // [capacity] x 8
    uint32_t* count;
    uint32_t n_lods, per_lod, capacity, n_tiles, tile_size, tiles_per_chunk, n_long, n_chunks;  // chunk_range
    uint32_t* err;                // [0] set when the scan gives up its bounded wait
    uint32_t spin_limit;
};
// The LOD-instanced output's device state: `hist`, the chunk x LOD histogram (n_chunks x n_lods words, every word
// written by the decide kernel or the histogram kernel), and `seg`, inst_seg_words() 64-bit words in which the scan
// publishes each segment of kInstSegChunks chunks' totals with a presence flag; zero between calls (the scatter
// launch clears them). 64 chunks per segment: at 767 chunks the scan took 5.7 us against 5.9 us with 32 and 7.1 us
// with 16, and the scatter clears half as many words.
constexpr uint32_t kInstSegChunks = 64;
inline uint32_t inst_segments(uint32_t n_chunks) { return n_chunks ? (n_chunks + kInstSegChunks - 1) / kInstSegChunks : 1u; }
inline uint64_t inst_seg_words(uint32_t n_chunks, uint32_t n_lods) { return (uint64_t)inst_segments(n_chunks) * n_lods; }
// LODs the decide kernel can histogram in LDS (more: a separate histogram kernel reads the decided entries).
constexpr uint32_t kInstLdsBins = 4096;
// with_hist: the decide kernel already wrote the histogram (chunks = its workgroups' tile ranges).
hipError_t launch_instanced(const InstancedJob& job, const CompactScratch& s, uint32_t* hist, unsigned long long* seg,
                            bool with_hist, hipStream_t st);
// The decide of the LOD-instanced output with its LDS histogram: the workgroup count and tiles per workgroup
// (= chunk) it will use for `job`, and the launch.
void decide_inst_chunks(const CullJob& job, uint32_t block, uint32_t items, uint32_t n_cu, uint32_t per_cu,
                        uint32_t* tiles_per_chunk, uint32_t* n_long, uint32_t* n_chunks);

// LATE pass from a frame cache recorded with the same tile size (block x items) over the same range. It runs over
// super-tiles of kSuperTile objects and needs late_cached_tiles(n) counts and late_cached_entries(n) staging slots.
constexpr uint32_t kSuperTile = 4096;
inline uint32_t late_cached_tiles(uint32_t n) { return (uint32_t)(((uint64_t)n + kSuperTile - 1) / kSuperTile); }
inline uint64_t late_cached_entries(uint32_t n) { return (uint64_t)late_cached_tiles(n) * kSuperTile; }
hipError_t launch_late_cached(const CullJob& job, const CompactScratch& s, const FrameCacheDev& cache, uint32_t block,
                              uint32_t items, hipStream_t st);
// The first tiles of up to kArgJobs jobs, passed to the batched decide as a kernel argument (2 KiB of the 4 KiB
// kernarg segment).
constexpr uint32_t kArgJobs = 512;
struct JobStartsArg {
    uint32_t v[kArgJobs];
};
// jobs_host: the same n_jobs jobs in host memory (their tile_start values fill the kernel-argument table)
// starts_dev: the jobs' tile_start values as a compact ascending array (the batched kernels stage it in LDS)
// max_job_tiles: the largest job's tile count (above kScanTiles the emit reads a scanned prefix of the batch)
hipError_t launch_cull_batch(const CullJob* jobs_host, const CullJob* jobs_dev, const uint32_t* starts_dev,
                             uint32_t n_jobs, uint32_t total_tiles, uint32_t max_job_tiles, const CompactScratch& s,
                             uint32_t block, uint32_t items, hipStream_t st);
inline uint32_t max_job_tiles(const CullJob* jobs, size_t n)
{
    uint32_t m = 0;
    for (size_t k = 0; k < n; ++k) m = jobs[k].n_tiles > m ? jobs[k].n_tiles : m;
    return m;
}
hipError_t launch_pyramid(const PyrBuildJob& job, uint32_t* err, hipStream_t st);
// starts_dev: the jobs' wg_start values as a compact ascending array (the batched kernels stage it in LDS)
// Many pyramid builds in one launch, one workgroup per (level-0 tile, job); max_tiles = the largest job's tiles.
hipError_t launch_pyramid_batch(const PyrBuildJob* jobs_dev, uint32_t n_jobs, uint32_t max_tiles, uint32_t* err,
                                hipStream_t st);
hipError_t launch_pyramid_level(const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw, uint32_t dh,
                                hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t st);
// Staged updates: ids[i] and the 32-byte MeshTransform record i (pos.xyz, scale | quaternion) scattered into the SoA
// streams, the last occurrence of an id winning; an id >= n_xf is skipped and sets err[2]. dev_ids: n device words
// (a copy of the ids); winner: n_xf device words, all zero before and after the launches.
hipError_t launch_scatter_staged_transforms(const uint32_t* ids, const float4* records, uint32_t n, uint32_t n_xf,
                                            uint32_t* dev_ids, uint32_t* winner, float4* xf_ps, float4* xf_q,
                                            uint32_t* err, hipStream_t st);
hipError_t launch_synth_depth(uint64_t seed, uint32_t w, uint32_t h, uint32_t occluders, float p11, float znear,
                              float* depth, hipStream_t st);

// Cluster two-level cull
// One synthetic depth buffer of a batched raster launch (one workgroup per 64 x 64 pixel tile). The first
// workgroup of a job also clears `n_tail` words at `tail_zero` (the flagged words of a pyramid built next).
struct SynthJob {
    uint64_t seed;
    float* depth;
    unsigned long long* tail_zero;
    uint32_t w, h, occluders, tiles_x;
    float p11, znear;
    uint32_t n_tail, wg_start;
};
hipError_t launch_synth_depth_batch(const SynthJob* jobs_dev, const uint32_t* starts_dev, uint32_t n_jobs,
                                    uint32_t total_wgs, hipStream_t st);
constexpr uint32_t kSynthTile = 64;       // raster tile edge
constexpr uint32_t kSynthMaxRects = 4096; // occluder rectangles per depth buffer

// Cluster pass of the two-level cluster cull (the object pass is a CullJob run by launch_cluster_objects).
struct ClusterJob {
    SceneDev scene;
    ViewConsts view;
    PyramidDesc pyr;
    const float4* cl_sphere;    // [n_clusters] center, radius
    const uint32_t* cl_cone;    // [n_clusters] int8 axis x, y, z and cutoff, packed low to high
    const uint2* cl_draw;       // [n_clusters] indexCount (= triangleCount*3), firstIndex (= dataOffset)
    uint32_t* dispatch;         // ClusterDispatchData[dispatch_capacity] as 3 dwords each
    uint32_t* dispatch_count;
    uint32_t* cmds;             // IndirectDrawData[capacity] as 6 dwords each
    uint32_t* count;
    uint32_t dispatch_capacity, capacity;
    uint32_t tests, hiz;        // BZ_CLUSTER_* and whether a pyramid is bound
    uint32_t n_tiles, tile_size;
    uint32_t scanned;           // more than kScanTiles tiles: record bases and survivor prefixes come from scans
};

// Scratch of the cluster pass: per object tile the cluster total (written by the object pass), the first record
// and the survivor count; survivors (objectId, clusterId) compacted inside each tile's record range.
struct ClusterScratch {
    uint32_t* ccounts;          // [capacity_tiles]
    uint32_t* rbase;            // [capacity_tiles + 1]
    uint32_t* scount;           // [capacity_tiles]
    uint32_t* sprefix;          // [capacity_tiles + 1] exclusive prefix of scount (scanned passes)
    uint2* survivors;           // [capacity_records]
    uint32_t capacity_tiles;
    uint64_t capacity_records;
};

hipError_t launch_cluster_objects(const CullJob& job, const CompactScratch& s, uint32_t* ccounts, uint32_t block,
                                  uint32_t items, hipStream_t st);
hipError_t launch_cluster_pass(const ClusterJob& job, const CompactScratch& cs, const ClusterScratch& xs,
                               hipStream_t st);

}  // namespace bz
