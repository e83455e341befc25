/*
 * bzcull.h — C ABI of the MI355X (gfx950) GPU-driven draw-cull library.
 *
 * The library replaces the compute part of Blitzen's Vulkan frame:
 *   DrawCullFirstPass      (src/Renderer/BlitzenVulkan/vulkanDraw.cpp:107-158, shader
 *                           src/Renderer/VulkanShaders/InitialDrawCull.comp.glsl:12-58)
 *   GenerateDepthPyramid   (vulkanDraw.cpp:554-622, shader DepthPyramidGeneration.comp.glsl:13-19)
 *   DrawCullOcclusionPass  (vulkanDraw.cpp:162-226, shader LateDrawCull.comp.glsl:13-71)
 * and produces the same DrawIndexedIndirect command buffer + draw count + per-object visibility buffer,
 * with the reference's byte layouts (src/Renderer/Resources/renderingResourcesTypes.h,
 * src/Renderer/BlitzenVulkan/vulkanData.h:429-433, src/Game/blitCamera.h:38-64).
 *
 * Conventions
 *   - Every entry point returns a bz_status (0 = ok, < 0 = error) and never throws or aborts.
 *   - Host arrays passed in are copied before the call returns; the caller keeps ownership.
 *   - Pointers named *_dev are device pointers (hipMalloc'd memory or bz_dev_alloc).
 *   - `stream` is a hipStream_t (NULL = the default stream). Launching calls are asynchronous.
 *   - Output order of commands is ascending objectId (deterministic), unlike the reference's atomicAdd order.
 *   - A context is not thread-safe, except bz_server_submit / bz_server_poll (multi-producer).
 */
#ifndef BZCULL_H
#define BZCULL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BZCULL_ABI_VERSION 2u

#if defined(__GNUC__)
#define BZ_API __attribute__((visibility("default")))
#else
#define BZ_API
#endif

typedef int32_t bz_status;
#define BZ_OK            0
#define BZ_E_INVALID    -1  /* bad argument, bad handle, size mismatch            */
#define BZ_E_OOM        -2  /* device or host allocation failed                   */
#define BZ_E_HIP        -3  /* a HIP runtime call failed                          */
#define BZ_E_CAPACITY   -4  /* an output buffer is too small                       */
#define BZ_E_BUDGET     -5  /* the request does not fit the server's HBM budget    */
#define BZ_E_TIMEOUT    -6  /* a deadline passed                                  */
#define BZ_E_NOT_READY  -7  /* result not available yet                           */
#define BZ_E_NOT_FOUND  -8  /* unknown ticket / handle                            */
#define BZ_E_STATE      -9  /* call not valid in the current state                */
#define BZ_E_NODEVICE  -10  /* no usable GPU                                      */
#define BZ_E_INTERNAL  -11  /* a kernel reported an internal error                */
#define BZ_E_CANCELLED -12  /* request cancelled                                  */

/* ---------------------------------------------------------------------------------------------
 * Reference byte layouts (static_assert'ed in the implementation).
 * ------------------------------------------------------------------------------------------- */

/* RenderObject — renderingResourcesTypes.h:155-159 == ShaderBuffers.glsl:118-124 (8 B) */
typedef struct bz_RenderObject {
    uint32_t transformId;
    uint32_t surfaceId;
} bz_RenderObject;

/* MeshTransform — renderingResourcesTypes.h:124-129 == ShaderBuffers.glsl:141-146 (32 B) */
typedef struct bz_MeshTransform {
    float pos[3];
    float scale;
    float orientation[4]; /* quaternion x, y, z, w (need not be unit) */
} bz_MeshTransform;

/* PrimitiveSurface — renderingResourcesTypes.h:104-116 == ShaderBuffers.glsl:52-66 (32 B) */
typedef struct bz_PrimitiveSurface {
    float center[3];
    float radius;
    uint32_t materialId;
    uint32_t lodOffset;
    uint32_t lodCount;
    uint32_t vertexOffset;
} bz_PrimitiveSurface;

/* LodData — renderingResourcesTypes.h:73-96 == CullingShaderData.glsl:75-92 (32 B) */
typedef struct bz_LodData {
    uint32_t indexCount;
    uint32_t firstIndex;
    uint32_t clusterOffset;
    uint32_t clusterCount;
    float error;
    uint32_t padding0, padding1, padding2;
} bz_LodData;

/* Cluster — renderingResourcesTypes.h:27-47 == ShaderBuffers.glsl:22-39 (32 B) */
typedef struct bz_Cluster {
    float center[3];
    float radius;
    int8_t coneAxisX, coneAxisY, coneAxisZ, coneCutoff;
    uint32_t dataOffset;
    uint8_t vertexCount, triangleCount, padding0, padding1;
    uint32_t padding2; /* alignas(16) tail */
} bz_Cluster;

/* IndirectDrawData — vulkanData.h:429-433 == ShaderBuffers.glsl:73-84 (24 B):
 * drawId then VkDrawIndexedIndirectCommand {indexCount, instanceCount, firstIndex, vertexOffset, firstInstance} */
typedef struct bz_IndirectDrawData {
    uint32_t objectId;
    uint32_t indexCount;
    uint32_t instanceCount;
    uint32_t firstIndex;
    int32_t vertexOffset;
    uint32_t firstInstance;
} bz_IndirectDrawData;

/* ClusterDispatchData — vulkanData.h:442-447 == CullingShaderData.glsl:118-123 (12 B) */
typedef struct bz_ClusterDispatchData {
    uint32_t objectId;
    uint32_t lodIndex;  /* absolute index into the LOD table */
    uint32_t clusterId; /* absolute index into the cluster table */
} bz_ClusterDispatchData;

/* CameraViewData — src/Game/blitCamera.h:38-64 == ShaderBuffers.glsl:173-194 (256 B, lodTarget at 180).
 * Matrices are column-major (BlitML mat4, blitMLTypes.h:117-146). */
typedef struct bz_CameraViewData {
    float view[16];
    float projectionView[16];
    float position[3];
    float frustumRight;
    float frustumLeft;
    float frustumTop;
    float frustumBottom;
    float proj0;
    float proj5;
    float zNear;
    float zFar;
    float pyramidWidth;
    float pyramidHeight;
    float lodTarget;
    uint8_t padding[72];
} bz_CameraViewData;

/* ---------------------------------------------------------------------------------------------
 * Context
 * ------------------------------------------------------------------------------------------- */
typedef struct bz_ctx bz_ctx;
typedef uint64_t bz_scene;   /* 0 = invalid */
typedef uint64_t bz_pyramid; /* 0 = none    */

typedef struct bz_config {
    int32_t device;        /* HIP device ordinal */
    uint32_t flags;        /* reserved, 0 */
    uint64_t reserved[6];
} bz_config;

/* Replaces VulkanRenderer::Init's device setup for the cull path (vulkanRenderer.h:20). */
BZ_API bz_status bz_create(const bz_config* cfg, bz_ctx** out);
BZ_API void bz_destroy(bz_ctx* ctx);
BZ_API const char* bz_status_str(bz_status s);
/* Text of the last error recorded on this context (empty if none). */
BZ_API const char* bz_last_error(const bz_ctx* ctx);
BZ_API uint32_t bz_abi_version(void);

/* ---------------------------------------------------------------------------------------------
 * Scene residency — replaces StaticBuffersInit / CopyStaticBufferDataToGPUBuffers
 * (vulkanRendererSetup.cpp:313-363, 365-647): uploads the host arrays once, keeps them resident in HBM
 * as structure-of-arrays streams, and zero-fills the visibility buffer (vulkanRendererSetup.cpp:349).
 * ------------------------------------------------------------------------------------------- */
#define BZ_MAX_OBJECTS 0x80000000u /* objects per scene (2^31): tile, slot and record offsets are 32-bit */
BZ_API bz_status bz_scene_create(bz_ctx* ctx,
                          const bz_RenderObject* objects, uint32_t n_objects,
                          const bz_MeshTransform* transforms, uint32_t n_transforms,
                          const bz_PrimitiveSurface* surfaces, uint32_t n_surfaces,
                          const bz_LodData* lods, uint32_t n_lods,
                          bz_scene* out);
/* Optional cluster table for the two-level cluster cull (renderingResourcesTypes.h:27-47). */
BZ_API bz_status bz_scene_set_clusters(bz_ctx* ctx, bz_scene scene, const bz_Cluster* clusters, uint32_t n_clusters);
BZ_API bz_status bz_scene_destroy(bz_ctx* ctx, bz_scene scene);
/* Replaces UpdateObjectTransform + UpdateBuffers (vulkanDraw.cpp:816-820, 46-60): scatters n transforms
 * (host arrays) into the resident transform streams, ordered on `stream`. Ids must be < n_transforms (checked: the
 * call fails with nothing applied); an id named more than once takes its last occurrence, as the reference's
 * sequential copies into its staging buffer do.
 * As in the reference (updates between frames), a cull reading the scene on another stream, including the serving
 * loop's, is not ordered with the update: order them with the streams, or update between frames. */
BZ_API bz_status bz_scene_update_transforms(bz_ctx* ctx, bz_scene scene, const uint32_t* transform_ids,
                                     const bz_MeshTransform* transforms, uint32_t n, void* stream);
/* Caller-filled staging for transform updates, as the reference's persistently mapped staging buffer
 * (vulkanRendererSetup.cpp:298-302) that UpdateObjectTransform writes into (vulkanDraw.cpp:816-820): returns pinned
 * host arrays of `capacity` transform ids and MeshTransforms (32 B records, the reference layout). The caller fills the
 * first n entries and calls bz_scene_commit_transforms: a kernel on `stream` reads them straight from the pinned slot
 * and scatters them (no host copy or conversion). Two slots are used in turn: the caller fills one while the other's
 * scatter runs; handing a slot out again waits for its previous scatter. An id named more than once in one commit
 * takes its last occurrence (entries are applied as if in order).
 * Calling the staging function again before committing returns the same slot (reallocated, without its contents,
 * when the new capacity is larger). Ids are checked on the device: an id >= n_transforms is skipped and
 * bz_check(stream) reports BZ_E_INVALID. */
BZ_API bz_status bz_scene_transform_staging(bz_ctx* ctx, bz_scene scene, uint32_t capacity, uint32_t** ids,
                                            bz_MeshTransform** transforms);
BZ_API bz_status bz_scene_commit_transforms(bz_ctx* ctx, bz_scene scene, uint32_t n, void* stream);
/* Overwrites the last-frame visibility buffer from a host array of n_objects uint32 (0 / non-zero). */
BZ_API bz_status bz_scene_set_visibility(bz_ctx* ctx, bz_scene scene, const uint32_t* visibility, uint32_t n, void* stream);
/* Copies the visibility buffer to a host array (synchronizes `stream`). */
BZ_API bz_status bz_scene_get_visibility(bz_ctx* ctx, bz_scene scene, uint32_t* visibility, uint32_t n, void* stream);
/* Copies the resident transforms [first, first + n) back to a host array in the reference layout (synchronizes
 * `stream`). */
BZ_API bz_status bz_scene_get_transforms(bz_ctx* ctx, bz_scene scene, uint32_t first, uint32_t n,
                                         bz_MeshTransform* transforms, void* stream);
/* Device pointer of the scene's visibility buffer (uint32[n_objects]). */
BZ_API bz_status bz_scene_visibility_ptr(bz_ctx* ctx, bz_scene scene, uint32_t** visibility_dev);
BZ_API bz_status bz_scene_info(bz_ctx* ctx, bz_scene scene, uint32_t* n_objects, uint32_t* n_transforms,
                        uint32_t* n_surfaces, uint32_t* n_lods, uint64_t* device_bytes);

/* ---------------------------------------------------------------------------------------------
 * Depth pyramid — replaces CreateDepthPyramid (vulkanResources.cpp:85-116) and GenerateDepthPyramid
 * (vulkanDraw.cpp:554-622). Level 0 samples the depth buffer through the MIN-reduction bilinear footprint
 * (vulkanResources.cpp:50-59); level i > 0 reduces level i-1 the same way. One kernel builds all levels.
 * Level i has extent max(1, base_w >> i) x max(1, base_h >> i), stored row-major, levels back to back.
 * ------------------------------------------------------------------------------------------- */
BZ_API bz_status bz_pyramid_create(bz_ctx* ctx, uint32_t base_w, uint32_t base_h, uint32_t mips, bz_pyramid* out);
BZ_API bz_status bz_pyramid_destroy(bz_ctx* ctx, bz_pyramid pyramid);
/* depth_dev: float[h][w] reverse-z depth (device). */
BZ_API bz_status bz_pyramid_build(bz_ctx* ctx, bz_pyramid pyramid, const float* depth_dev, uint32_t w, uint32_t h, void* stream);
/* Device pointer + total texel count of the pyramid storage, and the texel offset of each level. */
BZ_API bz_status bz_pyramid_storage(bz_ctx* ctx, bz_pyramid pyramid, float** texels_dev, uint64_t* n_texels,
                             uint32_t* level_offsets /* [mips] or NULL */);
/* The reference's extent rule: PreviousPow2 per axis and GetDepthPyramidMipLevels (blitML.h:54-62,95-107). */
BZ_API void bz_pyramid_reference_extent(uint32_t w, uint32_t h, uint32_t* base_w, uint32_t* base_h, uint32_t* mips);

/* ---------------------------------------------------------------------------------------------
 * Cull
 * ------------------------------------------------------------------------------------------- */
#define BZ_PASS_EARLY     0u /* InitialDrawCull.comp.glsl:12-58: objects with visibility != 0; frustum + LOD      */
#define BZ_PASS_LATE      1u /* LateDrawCull.comp.glsl:13-71: all objects; frustum + Hi-Z; emits newly visible;
                                rewrites visibility                                                                */
#define BZ_PASS_PLAIN     2u /* TransparentDrawCull.comp.glsl:13-65: all objects; frustum + LOD; no visibility     */
#define BZ_PASS_TEMPORAL  3u /* drawOccTemporal.hlsl:13-64 with the Vulkan Hi-Z sample: all objects; frustum +
                                Hi-Z against the given pyramid; emits all visible; no visibility                    */
#define BZ_PASS_ONPC      4u /* OnpcDrawCull.comp.glsl:11-52: as PLAIN but the LOD index is NOT offset by the
                                surface's lodOffset (the reference's behaviour, kept for parity)                   */

/* Flags */
#define BZ_CULL_EARLY_HIZ 0x1u /* EARLY pass also tests against `pyramid` (the previous frame's pyramid) */
/* Frame cache. EARLY: also runs the frustum test (and LOD selection) on every object of the range, last-frame
 * visible or not, and keeps the survivors' view-space spheres and LODs in the scene. LATE: when the scene holds
 * survivors recorded with the same view matrix, frustum constants, lodTarget and object range, and no transform
 * update since, reads those and the current visibility input (not every object's transform), and writes the
 * visibility of the range; otherwise runs in full. Outputs are identical either way, whatever was written to any
 * visibility buffer in between (by other passes, bz_scene_set_visibility or the caller's own kernels). Both calls go
 * on the same stream, or the caller orders them. */
#define BZ_CULL_FRAME_CACHE 0x2u
/* Frame-cache use of a scene: EARLY passes that recorded survivors, and LATE passes that read them. */
BZ_API bz_status bz_scene_frame_cache_stats(bz_ctx* ctx, bz_scene scene, uint64_t* recorded, uint64_t* reused);

typedef struct bz_cull_args {
    bz_scene scene;
    bz_pyramid pyramid;                 /* required for LATE and TEMPORAL, optional for EARLY (with BZ_CULL_EARLY_HIZ) */
    uint32_t pass;                      /* BZ_PASS_* */
    uint32_t flags;                     /* BZ_CULL_* */
    const bz_CameraViewData* view;      /* host pointer; copied at launch */
    bz_IndirectDrawData* cmds_dev;      /* output commands */
    uint32_t* count_dev;                /* output draw count (exact even when > capacity) */
    uint32_t capacity;                  /* commands beyond `capacity` are not written */
    uint32_t object_begin;              /* cull objects [object_begin, object_begin + object_count) */
    uint32_t object_count;              /* 0 with object_begin 0 = all objects */
    const uint32_t* visibility_in_dev;  /* NULL = the scene's visibility buffer */
    uint32_t* visibility_out_dev;       /* LATE only. NULL = visibility_in (in place, as the reference) */
} bz_cull_args;

/* One cull pass on `stream`. Replaces one vkCmdFillBuffer(count, 0) + vkCmdDispatch(N/64+1) of
 * DrawCullFirstPass / DrawCullOcclusionPass (vulkanDraw.cpp:119,145,180,213). Commands are written at
 * cmds_dev[0..min(count, capacity)); *count_dev = number of emitted commands. Overflow is detected by
 * the caller as *count_dev > capacity. */
BZ_API bz_status bz_cull(bz_ctx* ctx, const bz_cull_args* args, void* stream);

/* bz_cull without any device allocation. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream or graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, passes, outputs and status codes as bz_cull; the kernels' compaction scratch is a fixed device array of
 * 33 652 992 bytes that the library holds on every GPU it runs on until the process ends, instead of the per-stream
 * buffer bz_cull grows. At most BZ_NOALLOC_MAX_OBJECTS objects per call (BZ_E_INVALID otherwise, nothing launched).
 * BZ_CULL_FRAME_CACHE works once a bz_cull with the flag has allocated the scene's cache for the range (BZ_E_STATE
 * before). Nothing in the array needs to start at zero: every call rewrites what it reads. All calls on one GPU share
 * the array, so calls that overlap in time on one GPU (on different streams, or from several host threads) race;
 * calls ordered on one stream do not. */
#define BZ_NOALLOC_MAX_OBJECTS 4194304u
BZ_API bz_status bz_cull_noalloc(bz_ctx* ctx, const bz_cull_args* args, void* stream);

/* Cluster two-level cull (PreClusterDrawCull.comp.glsl:13-47 then InitialClusterCull.comp.glsl:12-55).
 * Object pass: frustum + LOD, as PreClusterDrawCull (with BZ_CLUSTER_OBJECT_HIZ in `flags`, also the Hi-Z test
 * against `pyramid`, which the reference's object pass does not have); expansion of the chosen LOD's clusters in
 * ascending (objectId, clusterId) order; cluster pass: per-cluster frustum, Hi-Z (against `pyramid`) and
 * backface-cone tests, as selected by `cluster_tests` (cluster tests are this library's definition; the reference
 * leaves them commented out). */
#define BZ_CLUSTER_FRUSTUM 0x1u
#define BZ_CLUSTER_HIZ     0x2u
#define BZ_CLUSTER_CONE    0x4u
#define BZ_CLUSTER_OBJECT_HIZ 0x1u /* bz_cluster_cull_args.flags */
typedef struct bz_cluster_cull_args {
    bz_scene scene;
    bz_pyramid pyramid;                       /* 0 = no Hi-Z test at either level */
    uint32_t cluster_tests;                   /* BZ_CLUSTER_FRUSTUM | BZ_CLUSTER_HIZ | BZ_CLUSTER_CONE */
    uint32_t flags;                           /* 0 or BZ_CLUSTER_OBJECT_HIZ */
    const bz_CameraViewData* view;
    bz_ClusterDispatchData* dispatch_dev;     /* expansion output (the reference's ClusterDispatchBuffer) */
    uint32_t* dispatch_count_dev;             /* number of expanded clusters */
    uint32_t dispatch_capacity;
    bz_IndirectDrawData* cmds_dev;            /* compacted cluster draws */
    uint32_t* count_dev;
    uint32_t capacity;
} bz_cluster_cull_args;
BZ_API bz_status bz_cull_clusters(bz_ctx* ctx, const bz_cluster_cull_args* args, void* stream);

/* ---------------------------------------------------------------------------------------------
 * LOD-instanced cull — replaces the D3D12 renderer's instanced path: drawInstCountReset.cs.hlsl,
 * drawInstCull.cs.hlsl and drawInstCmd.cs.hlsl, host DrawInstanceCullPass (dx12Draw.cpp:340-413).
 * Frustum + LOD per object (the BZ_PASS_PLAIN decision); each visible object joins the instance list of its
 * LOD; one instanced command per LOD with instances, in ascending LOD order. Deterministic: every list holds
 * its objects in ascending objectId; a list keeps at most instances_per_lod of them (the reference's
 * Ce_MaxInstanceCountPerLOD, blitzenEngine.h:65, overflows into the next list instead) while its counter keeps
 * the exact count, and the command draws min(count, instances_per_lod) instances.
 * ------------------------------------------------------------------------------------------- */
/* LodInstanceCounter — renderingResourcesTypes.h:98-102 */
typedef struct bz_LodInstanceCounter {
    uint32_t instanceOffset;   /* lod * instances_per_lod */
    uint32_t instanceCount;    /* exact number of visible objects at this LOD */
} bz_LodInstanceCounter;
/* IndirectDrawCmd — dx12Data.h:284-296 == cullBuffers.hlsl DrawCmd (32 B) */
typedef struct bz_IndirectDrawCmd32 {
    uint32_t objId;            /* the LOD's instanceOffset (root constant of the instanced draw) */
    uint32_t indexCount;
    uint32_t instanceCount;
    uint32_t firstIndex;
    int32_t vertexOffset;      /* 0 */
    uint32_t firstInstance;    /* 0 */
    uint32_t padding0, padding1;
} bz_IndirectDrawCmd32;
typedef struct bz_instanced_cull_args {
    bz_scene scene;
    const bz_CameraViewData* view;
    uint32_t instances_per_lod;
    uint32_t capacity;                        /* commands */
    uint32_t* instance_ids_dev;               /* [n_lods * instances_per_lod] object ids */
    bz_LodInstanceCounter* counters_dev;      /* [n_lods] */
    bz_IndirectDrawCmd32* cmds_dev;
    uint32_t* count_dev;                      /* number of commands (LODs with instances) */
} bz_instanced_cull_args;
BZ_API bz_status bz_cull_instanced(bz_ctx* ctx, const bz_instanced_cull_args* args, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Batched multi-view cull: many views (cameras, shadow cascades) of uneven object counts in one
 * persistent launch (late-pass semantics per view, each with its own pyramid, visibility and outputs).
 * ------------------------------------------------------------------------------------------- */
typedef struct bz_view_job {
    bz_scene scene;
    bz_pyramid pyramid;                 /* 0 = frustum + LOD only */
    uint32_t pass;                      /* BZ_PASS_LATE, BZ_PASS_PLAIN or BZ_PASS_TEMPORAL */
    uint32_t object_begin;
    uint32_t object_count;              /* 0 with object_begin 0 = all objects */
    uint32_t capacity;
    bz_CameraViewData view;             /* by value */
    const uint32_t* visibility_in_dev;  /* LATE: NULL = the scene's buffer */
    uint32_t* visibility_out_dev;       /* LATE: NULL = in place */
    bz_IndirectDrawData* cmds_dev;
    uint32_t* count_dev;
} bz_view_job;

typedef struct bz_tuning {
    uint32_t block;           /* threads per workgroup: 256 or 512 */
    uint32_t items;           /* objects per thread:    2 or 4     */
    uint32_t grid_per_cu;     /* workgroups per CU for the persistent decide kernel; 0 = as many as fit */
    uint32_t reserved;
} bz_tuning;

/* Culls every job in one launch pair (decide + emit over the jobs' concatenated tiles). Each job's output
 * equals what bz_cull gives for it alone. Jobs with no objects only get their count zeroed. The job table
 * is staged through a context-owned pinned buffer, so this call cannot be stream-captured: capture a fixed
 * batch with bz_batch_graph_create instead. Jobs that write one visibility buffer in overlapping ranges race. */
BZ_API bz_status bz_cull_batch(bz_ctx* ctx, const bz_view_job* jobs, uint32_t n_jobs, void* stream);
/* Synchronizes `stream` and reports BZ_E_INTERNAL if a kernel launched on it gave up a bounded wait
 * since the last check (the flag is then cleared). A pyramid build whose tail workgroup gave up waiting for a
 * tile's hand-off word can leave such words set by tiles that finished later: the check then also waits for the
 * device and clears the hand-off words of every pyramid of the context, so the next build of each is correct
 * (the pyramid that was being built holds wrong values until rebuilt). */
BZ_API bz_status bz_check(bz_ctx* ctx, void* stream);
/* The number of short sleeps (s_sleep 1, about 64 clocks each) a bounded wait inside a kernel may take before it
 * gives up; 0 gives up at the first word not yet handed over. Applies to launches made afterwards through this
 * context: pyramid builds (bz_pyramid_build, bz_pyramid_build_batch, bz_batch_graph_create) and the scan of
 * bz_cull_instanced. Default BZ_DEFAULT_WAIT_LIMIT.
 * A small limit forces the give-up path for tests of the error handling; with the default the waits end in a few
 * microseconds. */
#define BZ_DEFAULT_WAIT_LIMIT 1048576u
BZ_API bz_status bz_set_wait_limit(bz_ctx* ctx, uint32_t sleeps);
/* Tuning used by bz_cull (kernel 0) and bz_cull_batch (kernel 1); defaults are built in, bz_autotune
 * replaces them. */
BZ_API bz_status bz_get_tuning(bz_ctx* ctx, uint32_t kernel /* 0 = single view, 1 = batched */, bz_tuning* out);
BZ_API bz_status bz_set_tuning(bz_ctx* ctx, uint32_t kernel, const bz_tuning* t);
/* Times every candidate tuning (block 256/512 x items 2/4, and for kernel 0 a persistent grid of 0, 1, 2
 * or 4 workgroups per CU) on the given jobs with HIP events and keeps the fastest; results are identical for
 * every candidate. The jobs run repeatedly, so a LATE job should write its visibility out of place. If
 * cache_path is not NULL the choice is read from / written to that JSON file, keyed by device name, kernel
 * and floor(log2(total objects)); a cached entry is used without timing. */
BZ_API bz_status bz_autotune(bz_ctx* ctx, uint32_t kernel, const bz_view_job* jobs, uint32_t n_jobs,
                      const char* cache_path, bz_tuning* best, void* stream);

/* hipGraph capture of a fixed batch (pyramid builds + batched cull); replayed with bz_graph_launch. */
typedef uint64_t bz_graph;
typedef struct bz_pyramid_job {
    bz_pyramid pyramid;
    const float* depth_dev;
    uint32_t w, h;
} bz_pyramid_job;
BZ_API bz_status bz_batch_graph_create(bz_ctx* ctx, const bz_pyramid_job* pyr_jobs, uint32_t n_pyr,
                                const bz_view_job* jobs, uint32_t n_jobs, bz_graph* out);
BZ_API bz_status bz_graph_launch(bz_ctx* ctx, bz_graph graph, void* stream);
BZ_API bz_status bz_graph_destroy(bz_ctx* ctx, bz_graph graph);
/* Builds several pyramids in one launch. */
BZ_API bz_status bz_pyramid_build_batch(bz_ctx* ctx, const bz_pyramid_job* jobs, uint32_t n, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Camera constants — restates SetupCamera / UpdateProjection (src/Game/blitzenCamera.cpp:14-41, 93-126):
 * InfiniteZPerspective (blitML.h:261-278), frustum planes, lodTarget, view = Mat4Inverse(T * R(yaw, pitch)).
 * ------------------------------------------------------------------------------------------- */
BZ_API bz_status bz_camera_setup(float fov_radians, float window_w, float window_h, float z_near, float draw_distance,
                          const float position[3], float yaw, float pitch, bz_CameraViewData* out);

/* ---------------------------------------------------------------------------------------------
 * Serving: a request loop over resident scenes with continuous batching and HBM-budget admission.
 * ------------------------------------------------------------------------------------------- */
#define BZ_SERVER_RELEASE_ON_COMPLETE 0x1u /* free each request's commands at completion (count and times kept) */
typedef struct bz_server_cfg {
    uint64_t hbm_budget_bytes;     /* budget for per-request depth, pyramids and outputs */
    uint32_t max_inflight_batches; /* device batches in flight at once (streams) */
    uint32_t max_batch_requests;   /* upper bound of requests in one batch */
    uint32_t flags;                /* BZ_SERVER_* */
    uint32_t reserved0;
    uint64_t reserved[4];
} bz_server_cfg;

typedef struct bz_request {
    bz_scene scene;
    bz_CameraViewData view;        /* pyramidWidth/Height are filled by the server */
    uint64_t depth_seed;           /* the depth buffer is rasterised on the device from this seed */
    uint32_t depth_w, depth_h;
    uint32_t occluders;            /* occluder rectangles in the synthetic depth buffer */
    uint32_t pass;                 /* BZ_PASS_TEMPORAL (all visible) or BZ_PASS_PLAIN */
    uint64_t deadline_ns;          /* absolute CLOCK_MONOTONIC deadline; 0 = none */
    uint64_t user_tag;
} bz_request;

typedef struct bz_result {
    bz_status status;
    uint32_t count;                /* number of commands */
    uint64_t submit_ns;            /* CLOCK_MONOTONIC timestamps: entered the queue ... */
    uint64_t admit_ns;             /* ... admitted into a batch (budget reserved) ... */
    uint64_t complete_ns;          /* ... batch finished on the device */
    uint64_t batch_id;             /* which device batch served the request */
    uint32_t batch_size;           /* requests in that batch */
    uint32_t reserved;
    uint64_t user_tag;
    uint64_t arrival_ns;           /* when the request was due: its scheduled time for bz_server_submit_schedule
                                      (submit_ns - arrival_ns is the submitter's own delay), else submit_ns */
} bz_result;

BZ_API bz_status bz_server_start(bz_ctx* ctx, const bz_server_cfg* cfg);
BZ_API bz_status bz_server_stop(bz_ctx* ctx);
BZ_API bz_status bz_server_submit(bz_ctx* ctx, const bz_request* req, uint64_t* ticket);
/* Submits reqs[i] at offsets_ns[i] after the call starts (ascending offsets; a native busy-wait paces the
 * arrivals, e.g. a seeded Poisson process). tickets[i] = 0 and statuses[i] != BZ_OK for a rejected request;
 * statuses may be NULL. Returns when every request has been submitted. */
BZ_API bz_status bz_server_submit_schedule(bz_ctx* ctx, const bz_request* reqs, const uint64_t* offsets_ns, uint32_t n,
                                           uint64_t* tickets, bz_status* statuses);
/* Non-blocking. BZ_E_NOT_READY while the request is pending. */
BZ_API bz_status bz_server_poll(bz_ctx* ctx, uint64_t ticket, bz_result* out);
/* Blocks until done or timeout_ns elapses (BZ_E_TIMEOUT). */
BZ_API bz_status bz_server_wait(bz_ctx* ctx, uint64_t ticket, uint64_t timeout_ns, bz_result* out);
/* Copies a finished request's commands to a host array (up to capacity) and releases its buffers. */
BZ_API bz_status bz_server_fetch(bz_ctx* ctx, uint64_t ticket, bz_IndirectDrawData* cmds, uint32_t capacity, uint32_t* count);
BZ_API bz_status bz_server_cancel(bz_ctx* ctx, uint64_t ticket);
/* Releases a finished request's buffers without fetching. */
BZ_API bz_status bz_server_release(bz_ctx* ctx, uint64_t ticket);
typedef struct bz_server_stats {
    uint64_t submitted, completed, rejected, cancelled, batches;
    uint64_t budget_bytes, budget_in_use, budget_peak;
    uint32_t queue_depth, inflight_batches;
} bz_server_stats;
BZ_API bz_status bz_server_get_stats(bz_ctx* ctx, bz_server_stats* out);

/* ---------------------------------------------------------------------------------------------
 * Vulkan interop (Linux). The renderer keeps owning the cull outputs and the depth attachment as Vulkan memory
 * (the command + count buffers vkCmdDrawIndexedIndirectCount reads, vulkanDraw.cpp:470-471; the depth the pyramid
 * is built from, vulkanDraw.cpp:554-575) and exports them with VK_KHR_external_memory_fd (vkGetMemoryFdKHR, an
 * opaque fd: a dma-buf on amdgpu). The mapped device pointers go into bz_cull / bz_pyramid_build like any other.
 * Depth must be linear: copy the attachment into an exported buffer (vkCmdCopyImageToBuffer, D32_SFLOAT, tightly
 * packed rows = float[h][w]). Semaphores exported with VK_KHR_external_semaphore_fd order the two queues.
 * The caller's descriptor is never consumed: the library imports a duplicate, so the caller may close its own.
 * ------------------------------------------------------------------------------------------- */
typedef uint64_t bz_extmem;
typedef uint64_t bz_extsem;
/* Imports `size` bytes (the Vulkan allocation's size) of external memory from `fd`. */
BZ_API bz_status bz_extmem_import(bz_ctx* ctx, int fd, uint64_t size, bz_extmem* out);
/* Device pointer to bytes [offset, offset + size) of the import, valid until bz_extmem_release. */
BZ_API bz_status bz_extmem_map(bz_ctx* ctx, bz_extmem mem, uint64_t offset, uint64_t size, void** dev_ptr);
/* Synchronizes the device, unmaps every pointer of `mem` and releases the import. */
BZ_API bz_status bz_extmem_release(bz_ctx* ctx, bz_extmem mem);
#define BZ_EXTSEM_BINARY   0u /* VkSemaphore of VK_SEMAPHORE_TYPE_BINARY, opaque fd   */
#define BZ_EXTSEM_TIMELINE 1u /* VkSemaphore of VK_SEMAPHORE_TYPE_TIMELINE, opaque fd */
BZ_API bz_status bz_extsem_import(bz_ctx* ctx, int fd, uint32_t kind, bz_extsem* out);
/* Stream-ordered wait for / signal of the semaphore (`value` is the timeline value; ignored for binary). */
BZ_API bz_status bz_extsem_wait(bz_ctx* ctx, bz_extsem sem, uint64_t value, void* stream);
BZ_API bz_status bz_extsem_signal(bz_ctx* ctx, bz_extsem sem, uint64_t value, void* stream);
BZ_API bz_status bz_extsem_release(bz_ctx* ctx, bz_extsem sem);

/* ---------------------------------------------------------------------------------------------
 * Seeded synthetic workloads (SplitMix64 streams), identical on every host.
 * ------------------------------------------------------------------------------------------- */
typedef struct bz_synth_scene_cfg {
    uint64_t seed;
    uint32_t n_objects;
    uint32_t n_surfaces;
    uint32_t lods_per_surface;      /* index counts max(3, 3072 >> 2k), error 0.01 * 2^k */
    uint32_t clusters_min;          /* 0 = no clusters */
    uint32_t clusters_max;
    float pos_min, pos_max;         /* positions uniform in [pos_min, pos_max]^3 */
    float scale_min, scale_max;
} bz_synth_scene_cfg;
/* Fills caller-allocated host arrays (n_objects objects and transforms, n_surfaces surfaces,
 * n_surfaces * lods_per_surface LODs). */
BZ_API bz_status bz_synth_scene(const bz_synth_scene_cfg* cfg, bz_RenderObject* objects, bz_MeshTransform* transforms,
                         bz_PrimitiveSurface* surfaces, bz_LodData* lods);
/* Cluster table for a synthetic scene: returns the number of clusters when clusters == NULL. The LOD
 * table's clusterOffset/clusterCount are filled in (LOD 0 of each surface owns clusters, other LODs own
 * max(1, count >> 2k)). */
BZ_API bz_status bz_synth_clusters(const bz_synth_scene_cfg* cfg, bz_LodData* lods, bz_Cluster* clusters, uint32_t* n_clusters);
/* Synthetic reverse-z depth buffer (host): ground plane y = -10 seen from a camera with the given P11 and
 * zNear, plus `occluders` seeded rectangles of depth U[0.001, 0.2]; per pixel the nearest (max) depth. */
BZ_API bz_status bz_synth_depth(uint64_t seed, uint32_t w, uint32_t h, uint32_t occluders, float p11, float z_near, float* depth);
/* The same buffer rasterised on the device into depth_dev (float[h][w]). */
BZ_API bz_status bz_synth_depth_device(bz_ctx* ctx, uint64_t seed, uint32_t w, uint32_t h, uint32_t occluders, float p11,
                                float z_near, float* depth_dev, void* stream);
/* Seeded last-frame visibility: Bernoulli(p) per object. */
BZ_API bz_status bz_synth_visibility(uint64_t seed, uint32_t n, float p, uint32_t* visibility);
/* Seeded rotation updates of `count` distinct transforms (ids ascending). */
BZ_API bz_status bz_synth_transform_updates(uint64_t seed, uint64_t frame, const bz_MeshTransform* base, uint32_t n_transforms,
                                     uint32_t count, uint32_t* ids, bz_MeshTransform* out);

/* ---------------------------------------------------------------------------------------------
 * Device plumbing for callers without their own HIP allocator (tests, benches, bindings).
 * ------------------------------------------------------------------------------------------- */
BZ_API bz_status bz_dev_alloc(bz_ctx* ctx, uint64_t bytes, void** out);
BZ_API bz_status bz_dev_free(bz_ctx* ctx, void* p);
BZ_API bz_status bz_memcpy_h2d(bz_ctx* ctx, void* dst_dev, const void* src, uint64_t bytes, void* stream);
BZ_API bz_status bz_memcpy_d2h(bz_ctx* ctx, void* dst, const void* src_dev, uint64_t bytes, void* stream);
BZ_API bz_status bz_memcpy_d2d(bz_ctx* ctx, void* dst_dev, const void* src_dev, uint64_t bytes, void* stream);
BZ_API bz_status bz_memset_d32(bz_ctx* ctx, void* dst_dev, uint32_t value, uint64_t count, void* stream);
BZ_API bz_status bz_stream_create(bz_ctx* ctx, void** stream);
BZ_API bz_status bz_stream_destroy(bz_ctx* ctx, void* stream);
BZ_API bz_status bz_stream_sync(bz_ctx* ctx, void* stream);
BZ_API bz_status bz_device_info(bz_ctx* ctx, char* name, uint32_t name_len, uint32_t* n_cu, uint64_t* hbm_bytes);

#ifdef __cplusplus
}
#endif
#endif /* BZCULL_H */
