/*
 * bz_oracle.c — CPU oracle for the draw-cull path. TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. It restates, in plain fp32 C compiled with -ffp-contract=off and no
 * fast-math, the arithmetic of Blitzen's cull shaders (GLSL, which cannot be compiled or run here) and
 * of the host camera code (C++). A product is fused into the add or subtract that follows it (fmaf, one
 * rounding) exactly where an expression is written with fmaf: in orc_rotate_quat, orc_sphere_to_view,
 * orc_frustum and orc_project_sphere (GLSL lets a compiler
 * contract a*b+c in shaders that declare nothing `precise`, as the reference's do; the GPU kernels fuse at the
 * same points):
 *
 *   orc_rotate_quat       src/Renderer/VulkanShaderHeaders/ShaderBuffers.glsl:197-200
 *   orc_sphere_to_view    src/Renderer/VulkanShaderHeaders/CullingShaderData.glsl:44-46
 *   orc_frustum           CullingShaderData.glsl:36-58
 *   orc_lod_select        CullingShaderData.glsl:103-116
 *   orc_project_sphere    CullingShaderData.glsl:8-33
 *   orc_hiz_level         CullingShaderData.glsl:66 (floor(log2(.)) + the sampler's LOD clamp,
 *                         src/Renderer/BlitzenVulkan/vulkanResources.cpp:395-429)
 *   orc_hiz_sample        CullingShaderData.glsl:68 with the LINEAR + REDUCTION_MODE_MIN + CLAMP_TO_EDGE
 *                         sampler (vulkanResources.cpp:50-59)
 *   orc_occlusion         CullingShaderData.glsl:60-73
 *   orc_cull              InitialDrawCull.comp.glsl:12-58 (EARLY), LateDrawCull.comp.glsl:13-71 (LATE),
 *                         TransparentDrawCull.comp.glsl:13-65 (PLAIN), HlslShaders/CS/drawOccTemporal.hlsl:13-64
 *                         (TEMPORAL), OnpcDrawCull.comp.glsl:11-52 (ONPC)
 *   orc_pyramid_build     DepthPyramidGeneration.comp.glsl:13-19 + vulkanDraw.cpp:554-622
 *   orc_camera_setup      src/Game/blitzenCamera.cpp:14-41, 93-126; BlitML (blitML.h:54-62, 115-121, 219-223,
 *                         261-278, 312-372, 497-534, 559-570; blitMLTypes.h:143-193)
 *   orc_cluster_*         PreClusterDrawCull.comp.glsl:13-47, InitialClusterCull.comp.glsl:12-55, plus the
 *                         cluster tests this library defines (the reference has none).
 *   orc_cull_instanced    HlslShaders/CS/drawInstCull.cs.hlsl:12-42 + drawInstCmd.cs.hlsl:9-38 (LOD-instanced output in
 *                         a deterministic order; host dx12Draw.cpp:340-413)
 *
 * Points the reference leaves to fixed-function hardware are settled by explicit rules, shared with the
 * HIP kernels:
 *   - LOD level  L = clamp(floor(log2(m)), 0, mips-1) by exponent extraction (exact), m = max(w, h); m < 1 -> 0.
 *   - footprint  tx = u*W - 0.5 (two roundings), i0 = floor(tx), i1 = i0 + 1, both clamped to [0, W-1]; texel i1 counts only
 *                when its bilinear weight frac(tx) is non-zero (the MIN reduction ignores zero-weight texels).
 *   - output order ascending objectId (the reference's atomicAdd order is not deterministic).
 *
 * Pinning: tests/test_oracle_*.py check these functions against closed forms, brute force in float64,
 * golden vectors from the reference's compiled camera code (oracle/_ref) and invariants.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EXPORT __attribute__((visibility("default")))

/* ---- layouts (renderingResourcesTypes.h; vulkanData.h:429-433; blitCamera.h:38-64) ---- */
typedef struct { uint32_t transformId, surfaceId; } OrcObject;                                  /* 8 B  */
typedef struct { float pos[3]; float scale; float q[4]; } OrcTransform;                         /* 32 B */
typedef struct { float center[3]; float radius; uint32_t materialId, lodOffset, lodCount, vertexOffset; } OrcSurface; /* 32 B */
typedef struct { uint32_t indexCount, firstIndex, clusterOffset, clusterCount; float error; uint32_t pad[3]; } OrcLod; /* 32 B */
typedef struct { float center[3]; float radius; int8_t ax, ay, az, cutoff; uint32_t dataOffset;
                 uint8_t vertexCount, triangleCount, p0, p1; uint32_t p2; } OrcCluster;          /* 32 B */
typedef struct { uint32_t objectId, indexCount, instanceCount, firstIndex; int32_t vertexOffset; uint32_t firstInstance; } OrcCmd; /* 24 B */
typedef struct { uint32_t objectId, lodIndex, clusterId; } OrcDispatch;                        /* 12 B */
typedef struct {
    float view[16]; float projView[16]; float position[3];
    float frustumRight, frustumLeft, frustumTop, frustumBottom, proj0, proj5, zNear, zFar,
          pyramidWidth, pyramidHeight, lodTarget;
    uint8_t pad[72];
} OrcView;                                                                                      /* 256 B */

_Static_assert(sizeof(OrcObject) == 8, "RenderObject");
_Static_assert(sizeof(OrcTransform) == 32, "MeshTransform");
_Static_assert(sizeof(OrcSurface) == 32, "PrimitiveSurface");
_Static_assert(sizeof(OrcLod) == 32, "LodData");
_Static_assert(sizeof(OrcCluster) == 32, "Cluster");
_Static_assert(sizeof(OrcCmd) == 24, "IndirectDrawData");
_Static_assert(sizeof(OrcView) == 256, "CameraViewData");
_Static_assert(offsetof(OrcView, lodTarget) == 180, "lodTarget offset");

enum { ORC_EARLY = 0, ORC_LATE = 1, ORC_PLAIN = 2, ORC_TEMPORAL = 3, ORC_ONPC = 4 };
#define ORC_EARLY_HIZ 0x1u

typedef struct {
    const float* texels;   /* all levels back to back */
    uint32_t base_w, base_h, mips;
    uint32_t level_offset[32];
} OrcPyramid;

/* ================================================================================================
 * Per-object arithmetic
 * ============================================================================================== */

/* GLSL max(x, y) is "y if x < y else x"; kept literal so NaN propagation matches. */
static inline float glsl_max(float x, float y) { return (x < y) ? y : x; }
static inline float glsl_min(float x, float y) { return (y < x) ? y : x; }

/* ShaderBuffers.glsl:197-200: v + 2.0 * cross(q.xyz, cross(q.xyz, v) + q.w * v).
 * cross(a, b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x) (blitML.h:219-221). */
ORC_EXPORT void orc_rotate_quat(const float v[3], const float q[4], float out[3])
{
    float t0 = fmaf(q[3], v[0], fmaf(q[1], v[2], -(q[2] * v[1])));
    float t1 = fmaf(q[3], v[1], fmaf(q[2], v[0], -(q[0] * v[2])));
    float t2 = fmaf(q[3], v[2], fmaf(q[0], v[1], -(q[1] * v[0])));
    float c0 = fmaf(q[1], t2, -(q[2] * t1));
    float c1 = fmaf(q[2], t0, -(q[0] * t2));
    float c2 = fmaf(q[0], t1, -(q[1] * t0));
    out[0] = v[0] + 2.0f * c0;
    out[1] = v[1] + 2.0f * c1;
    out[2] = v[2] + 2.0f * c2;
}

/* CullingShaderData.glsl:44-46: center = RotateQuat(c, q) * scale + pos; center = (view * vec4(center, 1)).xyz;
 * radius = r * scale. mat*vec4 sums in the order of blitMLTypes.h:184-193 (m0*x + y*m4 + z*m8 + w*m12), unfused;
 * the scale-and-translate is fused. */
ORC_EXPORT void orc_sphere_to_view(const float c[3], float r, const float pos[3], float scale, const float q[4],
                                   const float* view, float out_c[3], float* out_r)
{
    float rq[3];
    orc_rotate_quat(c, q, rq);
    float w0 = fmaf(rq[0], scale, pos[0]);
    float w1 = fmaf(rq[1], scale, pos[1]);
    float w2 = fmaf(rq[2], scale, pos[2]);
    for (int k = 0; k < 3; ++k)
        out_c[k] = ((view[k] * w0 + w1 * view[4 + k]) + w2 * view[8 + k]) + 1.0f * view[12 + k];
    *out_r = r * scale;
}

/* CullingShaderData.glsl:52-55 (symmetric-frustum test on view-space sphere). */
ORC_EXPORT int orc_frustum(const float c[3], float r, float fRight, float fLeft, float fTop, float fBottom,
                           float zNear, float zFar)
{
    int vis = 1;
    vis = vis && (fmaf(c[2], fLeft, -(fabsf(c[0]) * fRight)) > -r);
    vis = vis && (fmaf(c[2], fBottom, -(fabsf(c[1]) * fTop)) > -r);
    vis = vis && (c[2] + r > zNear) && (c[2] - r < zFar);
    return vis;
}

/* CullingShaderData.glsl:103-116. Returns the index relative to lodOffset; the last i with error < threshold wins. */
ORC_EXPORT uint32_t orc_lod_select(const float c[3], float r, float scale, float lodTarget, uint32_t lodOffset,
                                   uint32_t lodCount, const OrcLod* lods)
{
    float len = sqrtf((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
    float distance = glsl_max(len - r, 0.0f);
    float threshold = (distance * lodTarget) / scale;
    uint32_t idx = 0;
    for (uint32_t i = 1; i < lodCount; ++i)
        if (lods[lodOffset + i].error < threshold) idx = i;
    return idx;
}

/* CullingShaderData.glsl:8-33 (Mara & McGuire 2013). aabb = (minU, minV, maxU, maxV) in UV space, v down. */
ORC_EXPORT int orc_project_sphere(const float c[3], float r, float znear, float P00, float P11, float aabb[4])
{
    if (c[2] < r + znear) return 0;
    float crx = c[0] * r, cry = c[1] * r, crz = c[2] * r;
    float czr2 = fmaf(c[2], c[2], -(r * r));
    float vx = sqrtf(fmaf(c[0], c[0], czr2));
    float minx = fmaf(vx, c[0], -crz) / fmaf(vx, c[2], crx);
    float maxx = fmaf(vx, c[0], crz) / fmaf(vx, c[2], -crx);
    float vy = sqrtf(fmaf(c[1], c[1], czr2));
    float miny = fmaf(vy, c[1], -crz) / fmaf(vy, c[2], cry);
    float maxy = fmaf(vy, c[1], crz) / fmaf(vy, c[2], -cry);
    /* vec4(minx*P00, miny*P11, maxx*P00, maxy*P11).xwzy * vec4(0.5, -0.5, 0.5, -0.5) + vec4(0.5) */
    aabb[0] = (minx * P00) * 0.5f + 0.5f;
    aabb[1] = (maxy * P11) * -0.5f + 0.5f;
    aabb[2] = (maxx * P00) * 0.5f + 0.5f;
    aabb[3] = (miny * P11) * -0.5f + 0.5f;
    return 1;
}

/* floor(log2(m)) clamped to the pyramid's mip range; exact for every float (exponent extraction). */
ORC_EXPORT uint32_t orc_hiz_level(float m, uint32_t mips)
{
    if (!(m >= 1.0f)) return 0; /* negative LOD and NaN clamp to level 0 */
    uint32_t bits;
    memcpy(&bits, &m, 4);
    uint32_t e = ((bits >> 23) & 0xffu) - 127u; /* m >= 1 is normal or inf: e in [0, 128] */
    return e < mips - 1 ? e : mips - 1;
}

static inline uint32_t clamp_texel(float f, uint32_t n)
{
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)(n - 1)) return n - 1;
    return (uint32_t)f;
}

/* Footprint of one sample (u, v) on a W x H level: up to 2 x 2 texels with non-zero bilinear weight. */
ORC_EXPORT float orc_footprint_min(const float* level, uint32_t W, uint32_t H, float u, float v)
{
    float tx = u * (float)W - 0.5f;
    float ty = v * (float)H - 0.5f;
    float fx = floorf(tx), fy = floorf(ty);
    uint32_t x0 = clamp_texel(fx, W), x1 = clamp_texel(fx + 1.0f, W);
    uint32_t y0 = clamp_texel(fy, H), y1 = clamp_texel(fy + 1.0f, H);
    int use_x1 = (tx - fx) > 0.0f;
    int use_y1 = (ty - fy) > 0.0f;
    float d = level[(size_t)y0 * W + x0];
    if (use_x1) d = glsl_min(d, level[(size_t)y0 * W + x1]);
    if (use_y1) {
        d = glsl_min(d, level[(size_t)y1 * W + x0]);
        if (use_x1) d = glsl_min(d, level[(size_t)y1 * W + x1]);
    }
    return d;
}

static inline uint32_t level_dim(uint32_t base, uint32_t L) { uint32_t d = base >> L; return d ? d : 1u; }

ORC_EXPORT float orc_hiz_sample(const OrcPyramid* p, uint32_t L, float u, float v)
{
    return orc_footprint_min(p->texels + p->level_offset[L], level_dim(p->base_w, L), level_dim(p->base_h, L), u, v);
}

/* CullingShaderData.glsl:60-73 with the level/footprint rules above. */
ORC_EXPORT int orc_occlusion(const float aabb[4], const OrcPyramid* p, float pyramidWidth, float pyramidHeight,
                             const float c[3], float r, float zNear)
{
    float width = (aabb[2] - aabb[0]) * pyramidWidth;
    float height = (aabb[3] - aabb[1]) * pyramidHeight;
    uint32_t L = orc_hiz_level(glsl_max(width, height), p->mips);
    float u = (aabb[0] + aabb[2]) * 0.5f;
    float v = (aabb[1] + aabb[3]) * 0.5f;
    float depth = orc_hiz_sample(p, L, u, v);
    float depthSphere = zNear / (c[2] - r);
    return depthSphere > depth;
}

/* ================================================================================================
 * Cull passes
 * ============================================================================================== */

typedef struct {
    const OrcObject* objects; uint32_t n_objects;
    const OrcTransform* transforms; uint32_t n_transforms;
    const OrcSurface* surfaces; uint32_t n_surfaces;
    const OrcLod* lods; uint32_t n_lods;
} OrcScene;

/* Decision for one object. Returns 1 when a command is emitted (lod_abs = absolute LOD index);
 * *vis_new = the visibility the LATE pass stores. */
static int cull_object(const OrcScene* s, const OrcView* v, const OrcPyramid* pyr, uint32_t pass, uint32_t flags,
                       uint32_t i, uint32_t vis_prev, uint32_t* lod_abs, uint32_t* vis_new)
{
    *vis_new = 0;
    if (pass == ORC_EARLY && vis_prev == 0) return 0; /* InitialDrawCull.comp.glsl:21-24 */
    const OrcObject o = s->objects[i];
    const OrcTransform* t = &s->transforms[o.transformId];
    const OrcSurface* sf = &s->surfaces[o.surfaceId];
    float c[3], r;
    orc_sphere_to_view(sf->center, sf->radius, t->pos, t->scale, t->q, v->view, c, &r);
    int visible = orc_frustum(c, r, v->frustumRight, v->frustumLeft, v->frustumTop, v->frustumBottom, v->zNear, v->zFar);
    int hiz = (pass == ORC_LATE || pass == ORC_TEMPORAL || (pass == ORC_EARLY && (flags & ORC_EARLY_HIZ)));
    if (hiz && pyr && visible) { /* LateDrawCull.comp.glsl:40-47 */
        float aabb[4];
        if (orc_project_sphere(c, r, v->zNear, v->proj0, v->proj5, aabb))
            visible = visible && orc_occlusion(aabb, pyr, v->pyramidWidth, v->pyramidHeight, c, r, v->zNear);
    }
    *vis_new = visible ? 1u : 0u;
    int emit = (pass == ORC_LATE) ? (visible && vis_prev == 0) : visible;
    if (!emit) return 0;
    uint32_t rel = orc_lod_select(c, r, t->scale, v->lodTarget, sf->lodOffset, sf->lodCount, s->lods);
    /* OnpcDrawCull.comp.glsl:38-40 uses the relative index directly (no += lodOffset). */
    *lod_abs = (pass == ORC_ONPC) ? rel : rel + sf->lodOffset;
    return 1;
}

static inline void make_cmd(OrcCmd* cmd, uint32_t objectId, const OrcLod* lod)
{
    cmd->objectId = objectId;
    cmd->indexCount = lod->indexCount;
    cmd->instanceCount = 1;
    cmd->firstIndex = lod->firstIndex;
    cmd->vertexOffset = 0;
    cmd->firstInstance = 0;
}

/* One pass over objects [begin, begin+count). Commands in ascending objectId, only the first `capacity`
 * written; returns the exact number emitted. vis_out (LATE) may alias vis_in. lod_out (optional) receives
 * the absolute LOD index of each emitted command. threads <= 0: sequential. */
ORC_EXPORT int64_t orc_cull(const OrcObject* objects, uint32_t n_objects, const OrcTransform* transforms,
                            uint32_t n_transforms, const OrcSurface* surfaces, uint32_t n_surfaces, const OrcLod* lods,
                            uint32_t n_lods, const OrcView* view, const OrcPyramid* pyr, uint32_t pass, uint32_t flags,
                            uint32_t begin, uint32_t count, const uint32_t* vis_in, uint32_t* vis_out, OrcCmd* cmds,
                            uint32_t capacity, uint32_t* lod_out, int threads)
{
    OrcScene s = { objects, n_objects, transforms, n_transforms, surfaces, n_surfaces, lods, n_lods };
    if (begin == 0 && count == 0) count = n_objects;
    if ((uint64_t)begin + count > n_objects) return -1;
    int nt = 1;
#ifdef _OPENMP
    if (threads > 0) nt = threads;
#endif
    if (nt > (int)count / 4096 + 1) nt = (int)count / 4096 + 1;
    uint64_t* chunk_count = (uint64_t*)calloc((size_t)nt + 1, sizeof(uint64_t));
    if (!chunk_count) return -2;
    /* pass 1: decisions per chunk (counts only) */
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int c = 0; c < nt; ++c) {
        uint32_t lo = begin + (uint32_t)(((uint64_t)count * c) / nt), hi = begin + (uint32_t)(((uint64_t)count * (c + 1)) / nt);
        uint64_t n = 0;
        for (uint32_t i = lo; i < hi; ++i) {
            uint32_t la, vn;
            uint32_t vp = vis_in ? vis_in[i] : 0;
            n += (uint64_t)cull_object(&s, view, pyr, pass, flags, i, vp, &la, &vn);
        }
        chunk_count[c + 1] = n;
    }
    for (int c = 0; c < nt; ++c) chunk_count[c + 1] += chunk_count[c];
    /* pass 2: emission at the chunk's prefix; visibility written after every read of the chunk */
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int c = 0; c < nt; ++c) {
        uint32_t lo = begin + (uint32_t)(((uint64_t)count * c) / nt), hi = begin + (uint32_t)(((uint64_t)count * (c + 1)) / nt);
        uint64_t pos = chunk_count[c];
        for (uint32_t i = lo; i < hi; ++i) {
            uint32_t la = 0, vn = 0;
            uint32_t vp = vis_in ? vis_in[i] : 0;
            int e = cull_object(&s, view, pyr, pass, flags, i, vp, &la, &vn);
            if (pass == ORC_LATE && vis_out) vis_out[i] = vn; /* LateDrawCull.comp.glsl:71 */
            if (e) {
                if (pos < capacity) {
                    make_cmd(&cmds[pos], i, &lods[la]);
                    if (lod_out) lod_out[pos] = la;
                }
                ++pos;
            }
        }
    }
    int64_t total = (int64_t)chunk_count[nt];
    free(chunk_count);
    return total;
}

/* LOD-instanced output (drawInstCountReset / drawInstCull / drawInstCmd .cs.hlsl; host dx12Draw.cpp:340-413) with
 * this library's deterministic rules: objects visit in ascending index; each visible object (the PLAIN decision,
 * as drawInstCull's frustum + LODSelection) is appended to its LOD's list while the list holds fewer than per_lod;
 * the counter counts every one. Commands, one per LOD with instances, in ascending LOD:
 * {lod * per_lod, indexCount, min(count, per_lod), firstIndex, 0, 0, 0, 0}. Returns the number of commands. */
ORC_EXPORT int64_t orc_cull_instanced(const OrcObject* objects, uint32_t n_objects, const OrcTransform* transforms,
                                      uint32_t n_transforms, const OrcSurface* surfaces, uint32_t n_surfaces,
                                      const OrcLod* lods, uint32_t n_lods, const OrcView* view, uint32_t per_lod,
                                      uint32_t* instance_ids, uint32_t* counters, uint32_t* cmds, uint32_t capacity)
{
    OrcScene s = { objects, n_objects, transforms, n_transforms, surfaces, n_surfaces, lods, n_lods };
    for (uint32_t l = 0; l < n_lods; ++l) {
        counters[2 * l] = l * per_lod;
        counters[2 * l + 1] = 0;
    }
    for (uint32_t i = 0; i < n_objects; ++i) {
        uint32_t la = 0, vn = 0;
        if (!cull_object(&s, view, NULL, ORC_PLAIN, 0, i, 0, &la, &vn)) continue;
        uint32_t k = counters[2 * la + 1]++;
        if (k < per_lod) instance_ids[(size_t)la * per_lod + k] = i;
    }
    int64_t n = 0;
    for (uint32_t l = 0; l < n_lods; ++l) {
        uint32_t c = counters[2 * l + 1];
        if (!c) continue;
        if ((uint64_t)n < capacity) {
            uint32_t* d = cmds + (size_t)n * 8;
            d[0] = l * per_lod;
            d[1] = lods[l].indexCount;
            d[2] = c < per_lod ? c : per_lod;
            d[3] = lods[l].firstIndex;
            d[4] = d[5] = d[6] = d[7] = 0;
        }
        ++n;
    }
    return n;
}

/* ================================================================================================
 * Depth pyramid (DepthPyramidGeneration.comp.glsl:13-19; host loop vulkanDraw.cpp:554-622)
 * dst(x, y) = MIN-footprint sample of src at ((x, y) + 0.5) / dstSize. src = depth for level 0, else level-1.
 * ============================================================================================== */
ORC_EXPORT void orc_pyramid_build(const float* depth, uint32_t w, uint32_t h, uint32_t base_w, uint32_t base_h,
                                  uint32_t mips, float* out, int threads)
{
    size_t off = 0, prev_off = 0;
    int nt = threads > 0 ? threads : 1;
    (void)nt;
    for (uint32_t L = 0; L < mips; ++L) {
        uint32_t dw = level_dim(base_w, L), dh = level_dim(base_h, L);
        const float* src = (L == 0) ? depth : out + prev_off;
        uint32_t sw = (L == 0) ? w : level_dim(base_w, L - 1), sh = (L == 0) ? h : level_dim(base_h, L - 1);
        float* dst = out + off;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (uint32_t y = 0; y < dh; ++y)
            for (uint32_t x = 0; x < dw; ++x) {
                float u = ((float)x + 0.5f) / (float)dw;
                float v = ((float)y + 0.5f) / (float)dh;
                dst[(size_t)y * dw + x] = orc_footprint_min(src, sw, sh, u, v);
            }
        prev_off = off;
        off += (size_t)dw * dh;
    }
}

ORC_EXPORT uint64_t orc_pyramid_texels(uint32_t base_w, uint32_t base_h, uint32_t mips, uint32_t* level_offsets)
{
    uint64_t off = 0;
    for (uint32_t L = 0; L < mips; ++L) {
        if (level_offsets) level_offsets[L] = (uint32_t)off;
        off += (uint64_t)level_dim(base_w, L) * level_dim(base_h, L);
    }
    return off;
}

ORC_EXPORT void orc_pyramid_init(OrcPyramid* p, const float* texels, uint32_t base_w, uint32_t base_h, uint32_t mips)
{
    p->texels = texels;
    p->base_w = base_w;
    p->base_h = base_h;
    p->mips = mips;
    orc_pyramid_texels(base_w, base_h, mips, p->level_offset);
}

/* ================================================================================================
 * Cluster two-level cull
 * ============================================================================================== */

/* Object pass (PreClusterDrawCull.comp.glsl:13-47): frustum (+ Hi-Z when pyr != NULL) + LOD, then one record per
 * cluster of the chosen LOD, in ascending (objectId, k). Returns the exact number of records. */
ORC_EXPORT int64_t orc_cluster_expand(const OrcObject* objects, uint32_t n_objects, const OrcTransform* transforms,
                                      uint32_t n_transforms, const OrcSurface* surfaces, uint32_t n_surfaces,
                                      const OrcLod* lods, uint32_t n_lods, const OrcView* view, const OrcPyramid* pyr,
                                      OrcDispatch* out, uint64_t capacity)
{
    OrcScene s = { objects, n_objects, transforms, n_transforms, surfaces, n_surfaces, lods, n_lods };
    uint64_t pos = 0;
    for (uint32_t i = 0; i < n_objects; ++i) {
        uint32_t la = 0, vn = 0;
        if (!cull_object(&s, view, pyr, pyr ? ORC_TEMPORAL : ORC_PLAIN, 0, i, 0, &la, &vn)) continue;
        const OrcLod* lod = &lods[la];
        for (uint32_t k = 0; k < lod->clusterCount; ++k, ++pos)
            if (pos < capacity) {
                out[pos].objectId = i;
                out[pos].lodIndex = la;
                out[pos].clusterId = lod->clusterOffset + k;
            }
    }
    return (int64_t)pos;
}

enum { ORC_CL_FRUSTUM = 1, ORC_CL_HIZ = 2, ORC_CL_CONE = 4 };

/* Cluster tests (this library's definition; InitialClusterCull.comp.glsl:24-42 leaves them commented out):
 *   sphere  : cluster sphere through the object transform and view (as orc_sphere_to_view)
 *   frustum : orc_frustum;  Hi-Z: orc_project_sphere + orc_occlusion
 *   cone    : axis = int8/127 rotated by the object quaternion and the view's 3x3, cutoff = int8/127;
 *             backfacing (culled) when dot(c, a) >= cutoff * length(c) + r (meshoptimizer's sphere form,
 *             camera at the view-space origin). */
ORC_EXPORT int orc_cluster_visible(const OrcCluster* cl, const OrcTransform* t, const OrcView* v, const OrcPyramid* pyr,
                                   uint32_t tests)
{
    float c[3], r;
    orc_sphere_to_view(cl->center, cl->radius, t->pos, t->scale, t->q, v->view, c, &r);
    int visible = 1;
    if (tests & ORC_CL_FRUSTUM)
        visible = orc_frustum(c, r, v->frustumRight, v->frustumLeft, v->frustumTop, v->frustumBottom, v->zNear, v->zFar);
    if (visible && (tests & ORC_CL_CONE)) {
        float a[3] = { (float)cl->ax / 127.0f, (float)cl->ay / 127.0f, (float)cl->az / 127.0f };
        float aw[3], av[3];
        orc_rotate_quat(a, t->q, aw);
        for (int k = 0; k < 3; ++k) av[k] = (v->view[k] * aw[0] + aw[1] * v->view[4 + k]) + aw[2] * v->view[8 + k];
        float cutoff = (float)cl->cutoff / 127.0f;
        float d = (c[0] * av[0] + c[1] * av[1]) + c[2] * av[2];
        float len = sqrtf((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
        if (d >= cutoff * len + r) visible = 0;
    }
    if (visible && (tests & ORC_CL_HIZ) && pyr) {
        float aabb[4];
        if (orc_project_sphere(c, r, v->zNear, v->proj0, v->proj5, aabb))
            visible = orc_occlusion(aabb, pyr, v->pyramidWidth, v->pyramidHeight, c, r, v->zNear);
    }
    return visible;
}

/* Cluster pass (InitialClusterCull.comp.glsl:12-55 + tests): one command per surviving record, in record order:
 * {objectId, triangleCount*3, 1, dataOffset, 0, 0}. Returns the exact count. */
ORC_EXPORT int64_t orc_cluster_cull(const OrcDispatch* recs, uint64_t n_recs, const OrcObject* objects,
                                    const OrcTransform* transforms, const OrcCluster* clusters, const OrcView* view,
                                    const OrcPyramid* pyr, uint32_t tests, OrcCmd* cmds, uint64_t capacity)
{
    uint64_t pos = 0;
    for (uint64_t k = 0; k < n_recs; ++k) {
        const OrcDispatch* d = &recs[k];
        const OrcCluster* cl = &clusters[d->clusterId];
        const OrcTransform* t = &transforms[objects[d->objectId].transformId];
        if (!orc_cluster_visible(cl, t, view, pyr, tests)) continue;
        if (pos < capacity) {
            cmds[pos].objectId = d->objectId;
            cmds[pos].indexCount = (uint32_t)cl->triangleCount * 3u;
            cmds[pos].instanceCount = 1;
            cmds[pos].firstIndex = cl->dataOffset;
            cmds[pos].vertexOffset = 0;
            cmds[pos].firstInstance = 0;
        }
        ++pos;
    }
    return (int64_t)pos;
}

/* ================================================================================================
 * Camera (src/Game/blitzenCamera.cpp:14-41, 93-126) — BlitML restated
 * ============================================================================================== */
typedef struct { float d[16]; } M4;

static M4 m4_identity(void) { M4 m; memset(&m, 0, sizeof m); m.d[0] = m.d[5] = m.d[10] = m.d[15] = 1.0f; return m; }

/* blitMLTypes.h:172-182: res[j + i*4] = a[j]*b[i*4] + a[4+j]*b[1+i*4] + a[8+j]*b[2+i*4] + a[12+j]*b[3+i*4] */
static M4 m4_mul(const M4* a, const M4* b)
{
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.d[j + i * 4] = ((a->d[0 + j] * b->d[0 + i * 4] + a->d[4 + j] * b->d[1 + i * 4]) + a->d[8 + j] * b->d[2 + i * 4])
                             + a->d[12 + j] * b->d[3 + i * 4];
    return r;
}

/* blitML.h:497-503 */
static void quat_normalize(const float q[4], float n[4])
{
    float len = sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) n[k] = q[k] / len;
}

/* blitML.h:520-534 */
ORC_EXPORT void orc_quat_to_mat4(const float q[4], float out[16])
{
    M4 r = m4_identity();
    float n[4];
    quat_normalize(q, n);
    float x = n[0], y = n[1], z = n[2], w = n[3];
    r.d[0] = (1.0f - 2.0f * y * y) - 2.0f * z * z;
    r.d[1] = 2.0f * x * y - 2.0f * z * w;
    r.d[2] = 2.0f * x * z + 2.0f * y * w;
    r.d[4] = 2.0f * x * y + 2.0f * z * w;
    r.d[5] = (1.0f - 2.0f * x * x) - 2.0f * z * z;
    r.d[6] = 2.0f * y * z - 2.0f * x * w;
    r.d[8] = 2.0f * x * z - 2.0f * y * w;
    r.d[9] = 2.0f * y * z + 2.0f * x * w;
    r.d[10] = (1.0f - 2.0f * x * x) - 2.0f * y * y;
    memcpy(out, r.d, sizeof r.d);
}

/* blitML.h:559-570 (normalize = 0) */
ORC_EXPORT void orc_quat_from_angle_axis(const float axis[3], float angle, float out[4])
{
    float half = 0.5f * angle;
    float s = sinf(half), c = cosf(half);
    out[0] = s * axis[0];
    out[1] = s * axis[1];
    out[2] = s * axis[2];
    out[3] = c;
}

/* blitML.h:312-363 */
ORC_EXPORT void orc_mat4_inverse(const float* m, float* out)
{
    float t0 = m[10] * m[15], t1 = m[14] * m[11], t2 = m[6] * m[15], t3 = m[14] * m[7];
    float t4 = m[6] * m[11], t5 = m[10] * m[7], t6 = m[2] * m[15], t7 = m[14] * m[3];
    float t8 = m[2] * m[11], t9 = m[10] * m[3], t10 = m[2] * m[7], t11 = m[6] * m[3];
    float t12 = m[8] * m[13], t13 = m[12] * m[9], t14 = m[4] * m[13], t15 = m[12] * m[5];
    float t16 = m[4] * m[9], t17 = m[8] * m[5], t18 = m[0] * m[13], t19 = m[12] * m[1];
    float t20 = m[0] * m[9], t21 = m[8] * m[1], t22 = m[0] * m[5], t23 = m[4] * m[1];
    float r[16];
    r[0] = ((t0 * m[5] + t3 * m[9]) + t4 * m[13]) - ((t1 * m[5] + t2 * m[9]) + t5 * m[13]);
    r[1] = ((t1 * m[1] + t6 * m[9]) + t9 * m[13]) - ((t0 * m[1] + t7 * m[9]) + t8 * m[13]);
    r[2] = ((t2 * m[1] + t7 * m[5]) + t10 * m[13]) - ((t3 * m[1] + t6 * m[5]) + t11 * m[13]);
    r[3] = ((t5 * m[1] + t8 * m[5]) + t11 * m[9]) - ((t4 * m[1] + t9 * m[5]) + t10 * m[9]);
    float d = 1.0f / (((m[0] * r[0] + m[4] * r[1]) + m[8] * r[2]) + m[12] * r[3]);
    r[0] = d * r[0];
    r[1] = d * r[1];
    r[2] = d * r[2];
    r[3] = d * r[3];
    r[4] = d * (((t1 * m[4] + t2 * m[8]) + t5 * m[12]) - ((t0 * m[4] + t3 * m[8]) + t4 * m[12]));
    r[5] = d * (((t0 * m[0] + t7 * m[8]) + t8 * m[12]) - ((t1 * m[0] + t6 * m[8]) + t9 * m[12]));
    r[6] = d * (((t3 * m[0] + t6 * m[4]) + t11 * m[12]) - ((t2 * m[0] + t7 * m[4]) + t10 * m[12]));
    r[7] = d * (((t4 * m[0] + t9 * m[4]) + t10 * m[8]) - ((t5 * m[0] + t8 * m[4]) + t11 * m[8]));
    r[8] = d * (((t12 * m[7] + t15 * m[11]) + t16 * m[15]) - ((t13 * m[7] + t14 * m[11]) + t17 * m[15]));
    r[9] = d * (((t13 * m[3] + t18 * m[11]) + t21 * m[15]) - ((t12 * m[3] + t19 * m[11]) + t20 * m[15]));
    r[10] = d * (((t14 * m[3] + t19 * m[7]) + t22 * m[15]) - ((t15 * m[3] + t18 * m[7]) + t23 * m[15]));
    r[11] = d * (((t17 * m[3] + t20 * m[7]) + t23 * m[11]) - ((t16 * m[3] + t21 * m[7]) + t22 * m[11]));
    r[12] = d * (((t14 * m[10] + t17 * m[14]) + t13 * m[6]) - ((t16 * m[14] + t12 * m[6]) + t15 * m[10]));
    r[13] = d * (((t20 * m[14] + t12 * m[2]) + t19 * m[10]) - ((t18 * m[10] + t21 * m[14]) + t13 * m[2]));
    r[14] = d * (((t18 * m[6] + t23 * m[14]) + t15 * m[2]) - ((t22 * m[14] + t14 * m[2]) + t19 * m[6]));
    r[15] = d * (((t22 * m[10] + t16 * m[2]) + t21 * m[6]) - ((t20 * m[6] + t23 * m[10]) + t17 * m[2]));
    memcpy(out, r, sizeof r);
}

/* blitML.h:223: p / Length(p.xyz), Length = sqrt(x*x + y*y + z*z) (blitML.h:115-121) */
static void normalize_plane(const float p[4], float out[4])
{
    float len = sqrtf((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
    for (int k = 0; k < 4; ++k) out[k] = p[k] / len;
}

/* SetupCamera (blitzenCamera.cpp:14-41) followed by UpdateProjection (:93-126). */
ORC_EXPORT void orc_camera_setup(float fov, float width, float height, float znear, float drawDistance,
                                 const float pos[3], float yaw, float pitch, OrcView* out)
{
    memset(out, 0, sizeof *out);
    out->zNear = znear;
    out->zFar = drawDistance;
    out->position[0] = pos[0];
    out->position[1] = pos[1];
    out->position[2] = pos[2];
    M4 T = m4_identity();
    T.d[12] = pos[0];
    T.d[13] = pos[1];
    T.d[14] = pos[2];
    const float yaw_axis[3] = { 0.0f, -1.0f, 0.0f }, pitch_axis[3] = { 1.0f, 0.0f, 0.0f };
    float qy[4], qp[4];
    orc_quat_from_angle_axis(yaw_axis, yaw, qy);
    orc_quat_from_angle_axis(pitch_axis, pitch, qp);
    M4 Ry, Rp;
    orc_quat_to_mat4(qy, Ry.d);
    orc_quat_to_mat4(qp, Rp.d);
    M4 R = m4_mul(&Ry, &Rp); /* blitzenCamera.cpp:6-12 */
    M4 TR = m4_mul(&T, &R);
    orc_mat4_inverse(TR.d, out->view);
    /* UpdateProjection */
    M4 P;
    memset(&P, 0, sizeof P);
    float halfTanFov = 1.0f / tanf(fov / 2.0f); /* blitML.h:261-278 */
    P.d[0] = halfTanFov / (width / height);
    P.d[5] = halfTanFov;
    P.d[11] = 1.0f;
    P.d[14] = znear;
    out->proj0 = P.d[0];
    out->proj5 = P.d[5];
    M4 V;
    memcpy(V.d, out->view, sizeof V.d);
    M4 PV = m4_mul(&P, &V);
    memcpy(out->projView, PV.d, sizeof PV.d);
    /* Transpose(P).GetRow(r) = (P[r], P[r+4], P[r+8], P[r+12]) */
    float r0[4] = { P.d[0], P.d[4], P.d[8], P.d[12] }, r1[4] = { P.d[1], P.d[5], P.d[9], P.d[13] };
    float r3[4] = { P.d[3], P.d[7], P.d[11], P.d[15] };
    float px[4], py[4], fx[4], fy[4];
    for (int k = 0; k < 4; ++k) { px[k] = r3[k] + r0[k]; py[k] = r3[k] + r1[k]; }
    normalize_plane(px, fx);
    normalize_plane(py, fy);
    out->frustumRight = fx[0];
    out->frustumLeft = fx[2];
    out->frustumTop = fy[1];
    out->frustumBottom = fy[2];
    out->lodTarget = (2.0f / out->proj5) * (1.0f / height);
}

ORC_EXPORT float orc_radians(float degrees) { return degrees * (3.14159265358979323846f / 180.0f); }

/* blitML.h:54-62 */
ORC_EXPORT uint32_t orc_previous_pow2(uint32_t value)
{
    uint32_t result = 1;
    while (result * 2 < value) result *= 2;
    return result;
}

/* blitML.h:95-107 */
ORC_EXPORT uint32_t orc_mip_levels(uint32_t width, uint32_t height)
{
    uint32_t n = 0;
    while (width > 1 || height > 1) { ++n; width /= 2; height /= 2; }
    return n;
}

ORC_EXPORT int orc_threads_available(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
