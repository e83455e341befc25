// synth.cpp — seeded synthetic workloads (host side).
//
// The reference builds its stress scene from unseeded rand() (src/Renderer/Resources/RenderObject/blitzenRender.cpp:
// 108-166, non-normalised quaternions). These generators are seeded and follow the workload description instead:
// positions uniform in a cube, scale uniform, unit quaternions (Shoemake's method from three uniforms), surfaces
// with bounding spheres (centre U[-1,1]^3, radius U[0.5,1.5]) and LOD chains whose index counts fall by 4x per level
// (max(3, 3072 >> 2k)) and whose error doubles per level (0.01 * 2^k), first indices cumulative.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bzcull.h"
#include "synth_rng.h"

using namespace bz;

extern "C" {

bz_status bz_synth_scene(const bz_synth_scene_cfg* cfg, bz_RenderObject* objects, bz_MeshTransform* transforms,
                         bz_PrimitiveSurface* surfaces, bz_LodData* lods)
{
    if (!cfg || cfg->n_surfaces == 0 || cfg->lods_per_surface == 0) return BZ_E_INVALID;
    if ((cfg->n_objects && (!objects || !transforms)) || !surfaces || !lods) return BZ_E_INVALID;
    const uint64_t seed = cfg->seed;
    const uint32_t S = cfg->n_surfaces, K = cfg->lods_per_surface;
    uint32_t first = 0;
    for (uint32_t s = 0; s < S; ++s) {
        bz_PrimitiveSurface& sf = surfaces[s];
        for (int k = 0; k < 3; ++k) sf.center[k] = uab(rng_at(seed, kStreamSurfSphere, (uint64_t)s * 4 + k), -1.0f, 1.0f);
        sf.radius = uab(rng_at(seed, kStreamSurfSphere, (uint64_t)s * 4 + 3), 0.5f, 1.5f);
        sf.materialId = 0;
        sf.lodOffset = s * K;
        sf.lodCount = K;
        sf.vertexOffset = 0;
        for (uint32_t k = 0; k < K; ++k) {
            bz_LodData& l = lods[s * K + k];
            const uint32_t shift = 2u * k;
            const uint32_t ic = shift < 32u ? (3072u >> shift) : 0u;
            l.indexCount = ic > 3u ? ic : 3u;
            l.firstIndex = first;
            first += l.indexCount;
            l.clusterOffset = 0;
            l.clusterCount = 0;
            l.error = 0.01f * (float)(1u << (k < 31u ? k : 31u));
            l.padding0 = l.padding1 = l.padding2 = 0;
        }
    }
    const float two_pi = 6.28318530717958647692f;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)cfg->n_objects; ++i) {
        bz_MeshTransform& t = transforms[i];
        for (int k = 0; k < 3; ++k)
            t.pos[k] = uab(rng_at(seed, kStreamPos, (uint64_t)i * 3 + k), cfg->pos_min, cfg->pos_max);
        t.scale = uab(rng_at(seed, kStreamScale, (uint64_t)i), cfg->scale_min, cfg->scale_max);
        const float u1 = u01(rng_at(seed, kStreamQuat, (uint64_t)i * 3 + 0));
        const float u2 = u01(rng_at(seed, kStreamQuat, (uint64_t)i * 3 + 1));
        const float u3 = u01(rng_at(seed, kStreamQuat, (uint64_t)i * 3 + 2));
        const float a = std::sqrt(1.0f - u1), b = std::sqrt(u1);
        t.orientation[0] = a * std::sin(two_pi * u2);
        t.orientation[1] = a * std::cos(two_pi * u2);
        t.orientation[2] = b * std::sin(two_pi * u3);
        t.orientation[3] = b * std::cos(two_pi * u3);
        objects[i].transformId = (uint32_t)i;
        objects[i].surfaceId = uint_below(rng_at(seed, kStreamSurface, (uint64_t)i), S);
    }
    return BZ_OK;
}

// Clusters: surface s owns C_s = U_int[clusters_min, clusters_max] clusters for LOD 0 and max(1, C_s >> 2k) for LOD k.
// Cluster sphere: centre = surface centre + U[-0.5,0.5]^3 * surface radius, radius U[0.2, 0.5] * surface radius.
// Normal cone: axis uniform on the unit sphere (z = U[-1,1], phi = U[0, 2pi)), cutoff U[-0.5, 0.9], both stored as
// 8-bit SNORM (round(x * 127)). 64 vertices and 64 triangles per cluster; dataOffset cumulative by 192 indices.
bz_status bz_synth_clusters(const bz_synth_scene_cfg* cfg, bz_LodData* lods, bz_Cluster* clusters, uint32_t* n_clusters)
{
    if (!cfg || !n_clusters || cfg->clusters_min == 0 || cfg->clusters_max < cfg->clusters_min) return BZ_E_INVALID;
    const uint64_t seed = cfg->seed;
    const uint32_t S = cfg->n_surfaces, K = cfg->lods_per_surface;
    uint64_t total = 0;
    std::vector<uint32_t> counts((size_t)S * K);
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t span = cfg->clusters_max - cfg->clusters_min + 1u;
        const uint32_t c0 = cfg->clusters_min + uint_below(rng_at(seed, kStreamClusters, s), span);
        for (uint32_t k = 0; k < K; ++k) {
            const uint32_t shift = 2u * k;
            const uint32_t c = shift < 32u ? (c0 >> shift) : 0u;
            counts[(size_t)s * K + k] = c ? c : 1u;
            total += counts[(size_t)s * K + k];
        }
    }
    if (total > 0xffffffffull) return BZ_E_INVALID;
    if (!clusters) {
        *n_clusters = (uint32_t)total;
        return BZ_OK;
    }
    if (!lods || *n_clusters < total) return BZ_E_CAPACITY;
    const float two_pi = 6.28318530717958647692f;
    uint32_t off = 0;
    for (uint32_t s = 0; s < S; ++s) {
        // surface sphere regenerated from the same stream as bz_synth_scene
        float sc[3], sr;
        for (int k = 0; k < 3; ++k) sc[k] = uab(rng_at(seed, kStreamSurfSphere, (uint64_t)s * 4 + k), -1.0f, 1.0f);
        sr = uab(rng_at(seed, kStreamSurfSphere, (uint64_t)s * 4 + 3), 0.5f, 1.5f);
        for (uint32_t k = 0; k < K; ++k) {
            bz_LodData& l = lods[s * K + k];
            l.clusterOffset = off;
            l.clusterCount = counts[(size_t)s * K + k];
            for (uint32_t c = 0; c < l.clusterCount; ++c, ++off) {
                bz_Cluster& cl = clusters[off];
                std::memset(&cl, 0, sizeof cl);
                const uint64_t b = (uint64_t)off * 8u;
                for (int a = 0; a < 3; ++a)
                    cl.center[a] = sc[a] + uab(rng_at(seed, kStreamClusterSphere, b + a), -0.5f, 0.5f) * sr;
                cl.radius = uab(rng_at(seed, kStreamClusterSphere, b + 3), 0.2f, 0.5f) * sr;
                const float z = uab(rng_at(seed, kStreamCone, b + 0), -1.0f, 1.0f);
                const float phi = two_pi * u01(rng_at(seed, kStreamCone, b + 1));
                const float rxy = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
                const float ax = rxy * std::cos(phi), ay = rxy * std::sin(phi);
                const float cut = uab(rng_at(seed, kStreamCone, b + 2), -0.5f, 0.9f);
                cl.coneAxisX = (int8_t)std::lrint(ax * 127.0f);
                cl.coneAxisY = (int8_t)std::lrint(ay * 127.0f);
                cl.coneAxisZ = (int8_t)std::lrint(z * 127.0f);
                cl.coneCutoff = (int8_t)std::lrint(cut * 127.0f);
                cl.dataOffset = off * 192u;
                cl.vertexCount = 64;
                cl.triangleCount = 64;
            }
        }
    }
    *n_clusters = off;
    return BZ_OK;
}

bz_status bz_synth_depth(uint64_t seed, uint32_t w, uint32_t h, uint32_t occluders, float p11, float z_near, float* depth)
{
    if (!depth || w == 0 || h == 0) return BZ_E_INVALID;
    for (uint32_t y = 0; y < h; ++y) {
        const float d = synth_plane_depth(y, h, p11, z_near);
        for (uint32_t x = 0; x < w; ++x) depth[(size_t)y * w + x] = d;
    }
    for (uint32_t r = 0; r < occluders; ++r) {
        const SynthRect s = synth_rect(seed, w, h, r);
        for (int32_t y = s.y0; y < s.y1; ++y)
            for (int32_t x = s.x0; x < s.x1; ++x) {
                float& d = depth[(size_t)y * w + x];
                d = (d < s.depth) ? s.depth : d;
            }
    }
    return BZ_OK;
}

bz_status bz_synth_visibility(uint64_t seed, uint32_t n, float p, uint32_t* visibility)
{
    if (n && !visibility) return BZ_E_INVALID;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) visibility[i] = u01(rng_at(seed, kStreamVis, (uint64_t)i)) < p ? 1u : 0u;
    return BZ_OK;
}

// `count` distinct ids chosen by a stride walk (ascending), each rotated by a seeded angle about a seeded axis:
// q' = q * (axis sin(a/2), cos(a/2)), renormalised.
bz_status bz_synth_transform_updates(uint64_t seed, uint64_t frame, const bz_MeshTransform* base, uint32_t n_transforms,
                                     uint32_t count, uint32_t* ids, bz_MeshTransform* out)
{
    if (!base || !ids || !out || count > n_transforms) return BZ_E_INVALID;
    if (count == 0) return BZ_OK;
    const uint64_t fseed = splitmix64(seed ^ (frame * 0x9E3779B97F4A7C15ull));
    const uint32_t stride = n_transforms / count;
    const uint32_t start = uint_below(rng_at(fseed, kStreamUpdate, 0), stride ? stride : 1u);
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t id = start + k * stride;
        ids[k] = id;
        const bz_MeshTransform& t = base[id];
        out[k] = t;
        const float ang = uab(rng_at(fseed, kStreamUpdate, 1 + (uint64_t)k * 4), -0.1f, 0.1f);
        float ax = uab(rng_at(fseed, kStreamUpdate, 2 + (uint64_t)k * 4), -1.0f, 1.0f);
        float ay = uab(rng_at(fseed, kStreamUpdate, 3 + (uint64_t)k * 4), -1.0f, 1.0f);
        float az = uab(rng_at(fseed, kStreamUpdate, 4 + (uint64_t)k * 4), -1.0f, 1.0f);
        const float al = std::sqrt(ax * ax + ay * ay + az * az) + 1e-6f;
        const float s = std::sin(0.5f * ang) / al, c = std::cos(0.5f * ang);
        const float rx = ax * s, ry = ay * s, rz = az * s, rw = c;
        const float* q = t.orientation;
        float nq[4] = { q[3] * rx + q[0] * rw + q[1] * rz - q[2] * ry, q[3] * ry - q[0] * rz + q[1] * rw + q[2] * rx,
                        q[3] * rz + q[0] * ry - q[1] * rx + q[2] * rw, q[3] * rw - q[0] * rx - q[1] * ry - q[2] * rz };
        const float nl = std::sqrt(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
        for (int j = 0; j < 4; ++j) out[k].orientation[j] = nq[j] / nl;
    }
    return BZ_OK;
}

}  // extern "C"
