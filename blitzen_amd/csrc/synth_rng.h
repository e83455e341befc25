// synth_rng.h — counter-based SplitMix64 streams shared by the host generators and the device depth rasteriser,
// so a seed gives bit-identical workloads on both sides.
// SplitMix64 (Steele, Lea, Flood 2014): x += 0x9E3779B97F4A7C15; z = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9;
// z = (z ^ (z >> 27)) * 0x94D049BB133111EB; return z ^ (z >> 31).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BZ_RNG_HD __host__ __device__ __forceinline__
#else
#define BZ_RNG_HD static inline
#endif

namespace bz {

BZ_RNG_HD uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Value `index` of stream `stream` under `seed`.
BZ_RNG_HD uint64_t rng_at(uint64_t seed, uint64_t stream, uint64_t index)
{
    const uint64_t key = splitmix64(seed ^ splitmix64(stream * 0xD1B54A32D192ED03ull));
    return splitmix64(key + index * 0x9E3779B97F4A7C15ull);
}

// Uniform float in [0, 1): 24 random bits.
BZ_RNG_HD float u01(uint64_t r) { return (float)(uint32_t)(r >> 40) * (1.0f / 16777216.0f); }

// Uniform float in [a, b): a + (b - a) * u (two fp32 roundings, no FMA).
BZ_RNG_HD float uab(uint64_t r, float a, float b)
{
    const float span = b - a;
    const float t = span * u01(r);
    return a + t;
}

// Uniform integer in [0, n) (multiply-shift, n < 2^32).
BZ_RNG_HD uint32_t uint_below(uint64_t r, uint32_t n) { return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32); }

// Stream ids
enum : uint64_t {
    kStreamPos = 1, kStreamScale = 2, kStreamQuat = 3, kStreamSurface = 4, kStreamSurfSphere = 5,
    kStreamRect = 6, kStreamVis = 7, kStreamClusters = 8, kStreamClusterSphere = 9, kStreamCone = 10,
    kStreamUpdate = 11
};

// One occluder rectangle of a synthetic depth buffer (inclusive-exclusive pixel bounds).
struct SynthRect {
    int32_t x0, y0, x1, y1;
    float depth;
};

// Rectangle r: width U_int[max(1, W/64), max(1, W/8)], height likewise from H, top-left uniform over the image
// (clipped at the right/bottom edge), depth U[0.001, 0.2].
BZ_RNG_HD SynthRect synth_rect(uint64_t seed, uint32_t W, uint32_t H, uint32_t r)
{
    const uint64_t b = (uint64_t)r * 8u;
    const uint32_t wmin = W / 64u ? W / 64u : 1u, wmax = W / 8u ? W / 8u : 1u;
    const uint32_t hmin = H / 64u ? H / 64u : 1u, hmax = H / 8u ? H / 8u : 1u;
    SynthRect s;
    const uint32_t rw = wmin + uint_below(rng_at(seed, kStreamRect, b + 0), wmax - wmin + 1u);
    const uint32_t rh = hmin + uint_below(rng_at(seed, kStreamRect, b + 1), hmax - hmin + 1u);
    s.x0 = (int32_t)uint_below(rng_at(seed, kStreamRect, b + 2), W);
    s.y0 = (int32_t)uint_below(rng_at(seed, kStreamRect, b + 3), H);
    s.x1 = s.x0 + (int32_t)rw < (int32_t)W ? s.x0 + (int32_t)rw : (int32_t)W;
    s.y1 = s.y0 + (int32_t)rh < (int32_t)H ? s.y0 + (int32_t)rh : (int32_t)H;
    s.depth = uab(rng_at(seed, kStreamRect, b + 4), 0.001f, 0.2f);
    return s;
}

// Ground plane y = -10 seen from a camera at the origin looking +z (reverse-z, infinite projection):
// pixel row py (row 0 at the top) has ndc y = 1 - 2 (py + 0.5) / H; below the horizon the plane's view depth is
// z = -10 P11 / ndc_y and the stored depth zNear / z = zNear * (-ndc_y) / (10 P11); at or above it, 0 (infinity).
BZ_RNG_HD float synth_plane_depth(uint32_t py, uint32_t H, float p11, float znear)
{
    const float t = (float)py + 0.5f;
    const float ndc = 1.0f - (2.0f * t) / (float)H;
    if (!(ndc < 0.0f)) return 0.0f;
    const float num = znear * (-ndc);
    const float den = 10.0f * p11;
    return num / den;
}

}  // namespace bz
