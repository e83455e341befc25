// cull_math.h — per-object cull arithmetic for the gfx950 kernels.
//
// fp32 only, compiled with -ffp-contract=off (no contraction by the compiler), IEEE denormals, correctly rounded
// sqrt and division, so every result is bit-identical to a CPU build of the same expressions. A product is fused
// into the add or subtract that follows it (one rounding: v_fma_f32) exactly where an expression is written with
// fma_f: in rotate_quat, sphere_to_view, frustum_visible and project_sphere. GLSL lets a compiler
// contract a*b+c in the reference's shaders (they declare nothing `precise`), so both forms are the reference's
// arithmetic; this one is fixed here and in the oracle, point for point, and takes about 14 fewer VALU instructions per
// object in the transform and frustum test.
// Sources of the arithmetic (GLSL, Blitzen):
//   rotate_quat        src/Renderer/VulkanShaderHeaders/ShaderBuffers.glsl:197-200
//   sphere_to_view     src/Renderer/VulkanShaderHeaders/CullingShaderData.glsl:44-46
//   frustum_visible    CullingShaderData.glsl:52-55
//   project_sphere     CullingShaderData.glsl:8-33
//   occlusion_visible  CullingShaderData.glsl:60-73 (+ sampler rules below)
//   lod_select         CullingShaderData.glsl:103-116
// The hardware-defined parts of the Hi-Z lookup are fixed here:
//   level     L = clamp(floor(log2(max(w, h))), 0, mips-1), floor(log2) by exponent extraction, m < 1 -> 0
//   footprint tx = u*W - 0.5 (two roundings); i0 = floor(tx), i1 = i0 + 1, clamped to [0, W-1] (CLAMP_TO_EDGE);
//             texel i1 is part of the MIN only when its weight frac(tx) is non-zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BZ_HD __host__ __device__ __forceinline__

namespace bz {

// Per-view constants used by the cull (a subset of CameraViewData, blitCamera.h:38-64).
struct ViewConsts {
    float m[12];      // view matrix rows 0..2: m[4*r + c] = view[c*4 + r] (column-major source), c = 0..3
    float fR, fL, fT, fB;
    float p00, p11;
    float znear, zfar;
    float pyrW, pyrH;
    float lodTarget;
    float pad;
};

// Device view of one depth pyramid.
struct PyramidDesc {
    const float* texels;
    uint32_t base_w, base_h, mips, pad;
    uint32_t level_offset[16];
};

// a * b + c with one rounding
BZ_HD float fma_f(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

BZ_HD float glsl_max(float x, float y) { return (x < y) ? y : x; }
BZ_HD float glsl_min(float x, float y) { return (y < x) ? y : x; }

BZ_HD void rotate_quat(float vx, float vy, float vz, float qx, float qy, float qz, float qw,
                       float& ox, float& oy, float& oz)
{
    // t = cross(q.xyz, v) + q.w * v ; r = v + 2 * cross(q.xyz, t); each cross term a*b - c*d as fma(a, b, -(c*d))
    const float tx = fma_f(qw, vx, fma_f(qy, vz, -(qz * vy)));
    const float ty = fma_f(qw, vy, fma_f(qz, vx, -(qx * vz)));
    const float tz = fma_f(qw, vz, fma_f(qx, vy, -(qy * vx)));
    ox = vx + 2.0f * fma_f(qy, tz, -(qz * ty));
    oy = vy + 2.0f * fma_f(qz, tx, -(qx * tz));
    oz = vz + 2.0f * fma_f(qx, ty, -(qy * tx));
}

// Bounding sphere (surface centre/radius) -> view space, through the object transform.
BZ_HD void sphere_to_view(const ViewConsts& v, float cx, float cy, float cz, float r, float px, float py, float pz,
                          float scale, float qx, float qy, float qz, float qw, float& ox, float& oy, float& oz,
                          float& orad)
{
    float rx, ry, rz;
    rotate_quat(cx, cy, cz, qx, qy, qz, qw, rx, ry, rz);
    const float wx = fma_f(rx, scale, px);
    const float wy = fma_f(ry, scale, py);
    const float wz = fma_f(rz, scale, pz);
    // column-major view * vec4(w, 1): ((m0*x + m4*y) + m8*z) + m12, unfused (the reference's mat * vec order)
    ox = ((v.m[0] * wx + v.m[1] * wy) + v.m[2] * wz) + v.m[3];
    oy = ((v.m[4] * wx + v.m[5] * wy) + v.m[6] * wz) + v.m[7];
    oz = ((v.m[8] * wx + v.m[9] * wy) + v.m[10] * wz) + v.m[11];
    orad = r * scale;
}

BZ_HD bool frustum_visible(const ViewConsts& v, float cx, float cy, float cz, float r)
{
    bool vis = fma_f(cz, v.fL, -(fabsf(cx) * v.fR)) > -r;
    vis = vis && (fma_f(cz, v.fB, -(fabsf(cy) * v.fT)) > -r);
    vis = vis && (cz + r > v.znear) && (cz - r < v.zfar);
    return vis;
}

// Returns false when the sphere crosses the near plane (no occlusion test then).
BZ_HD bool project_sphere(float cx, float cy, float cz, float r, float znear, float p00, float p11, float& a0,
                          float& a1, float& a2, float& a3)
{
    if (cz < r + znear) return false;
    const float crx = cx * r, cry = cy * r, crz = cz * r;
    const float czr2 = fma_f(cz, cz, -(r * r));
    const float vx = sqrtf(fma_f(cx, cx, czr2));
    const float minx = fma_f(vx, cx, -crz) / fma_f(vx, cz, crx);
    const float maxx = fma_f(vx, cx, crz) / fma_f(vx, cz, -crx);
    const float vy = sqrtf(fma_f(cy, cy, czr2));
    const float miny = fma_f(vy, cy, -crz) / fma_f(vy, cz, cry);
    const float maxy = fma_f(vy, cy, crz) / fma_f(vy, cz, -cry);
    a0 = (minx * p00) * 0.5f + 0.5f;
    a1 = (maxy * p11) * -0.5f + 0.5f;
    a2 = (maxx * p00) * 0.5f + 0.5f;
    a3 = (miny * p11) * -0.5f + 0.5f;
    return true;
}

BZ_HD uint32_t float_bits(float f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    uint32_t b;
    __builtin_memcpy(&b, &f, 4);
    return b;
#endif
}

BZ_HD uint32_t hiz_level(float m, uint32_t mips)
{
    if (!(m >= 1.0f)) return 0u;
    const uint32_t e = ((float_bits(m) >> 23) & 0xffu) - 127u;
    return e < mips - 1u ? e : mips - 1u;
}

BZ_HD uint32_t clamp_texel(float f, uint32_t n)
{
    if (!(f >= 0.0f)) return 0u;
    if (f >= (float)(n - 1u)) return n - 1u;
    return (uint32_t)f;
}

// Footprint coordinates of a sample; x1/y1 equal x0/y0 when that texel has zero weight.
struct Footprint {
    uint32_t x0, x1, y0, y1;
};

BZ_HD Footprint footprint(float u, float v, uint32_t W, uint32_t H)
{
    const float tx = u * (float)W - 0.5f;
    const float ty = v * (float)H - 0.5f;
    const float fx = floorf(tx), fy = floorf(ty);
    Footprint f;
    f.x0 = clamp_texel(fx, W);
    f.y0 = clamp_texel(fy, H);
    f.x1 = ((tx - fx) > 0.0f) ? clamp_texel(fx + 1.0f, W) : f.x0;
    f.y1 = ((ty - fy) > 0.0f) ? clamp_texel(fy + 1.0f, H) : f.y0;
    return f;
}

BZ_HD uint32_t level_dim(uint32_t base, uint32_t L)
{
    const uint32_t d = base >> L;
    return d ? d : 1u;
}

#if defined(__HIPCC__)
__device__ __forceinline__ float footprint_min(const float* __restrict__ level, uint32_t W, const Footprint& f)
{
    // min is exact and order-free; duplicates (zero-weight texels) do not change it.
    const float d00 = level[(size_t)f.y0 * W + f.x0];
    const float d01 = level[(size_t)f.y0 * W + f.x1];
    const float d10 = level[(size_t)f.y1 * W + f.x0];
    const float d11 = level[(size_t)f.y1 * W + f.x1];
    return glsl_min(glsl_min(d00, d01), glsl_min(d10, d11));
}

// The four footprint texels of a pyramid level as two 8-byte loads, one per row: texels (xa, xa + 1) with xa = x0,
// or x0 - 1 when x0 is the row's last texel, so that x0 and x1 (each x0 or x0 + 1, clamped) are both in the pair.
// On a level one texel wide the pair's second word is the texel after the row (the next row or level, or the spare
// texel every pyramid allocation holds past its last level) and is not used. The values are the ones four single
// loads return; the gathers take half the address-unit cycles.
__device__ __forceinline__ void footprint_texels(const float* __restrict__ level, uint32_t W, const Footprint& f,
                                                 float& d00, float& d01, float& d10, float& d11)
{
    const uint32_t xa = (f.x0 + 1u < W || f.x0 == 0u) ? f.x0 : f.x0 - 1u;
    float2 r0, r1;
    __builtin_memcpy(&r0, level + (size_t)f.y0 * W + xa, 8);
    __builtin_memcpy(&r1, level + (size_t)f.y1 * W + xa, 8);
    d00 = f.x0 == xa ? r0.x : r0.y;
    d01 = f.x1 == xa ? r0.x : r0.y;
    d10 = f.x0 == xa ? r1.x : r1.y;
    d11 = f.x1 == xa ? r1.x : r1.y;
}

// level_offset: the pyramid's per-level texel offsets (kept in LDS by the caller).
__device__ __forceinline__ bool occlusion_visible(const ViewConsts& v, const PyramidDesc& p, const uint32_t* level_offset,
                                                  float a0, float a1, float a2, float a3, float cz, float r)
{
    const float width = (a2 - a0) * v.pyrW;
    const float height = (a3 - a1) * v.pyrH;
    const uint32_t L = hiz_level(glsl_max(width, height), p.mips);
    const float u = (a0 + a2) * 0.5f;
    const float w = (a1 + a3) * 0.5f;
    const uint32_t W = level_dim(p.base_w, L), H = level_dim(p.base_h, L);
    const Footprint f = footprint(u, w, W, H);
    float d00, d01, d10, d11;
    footprint_texels(p.texels + level_offset[L], W, f, d00, d01, d10, d11);
    const float depth = glsl_min(glsl_min(d00, d01), glsl_min(d10, d11));
    const float depthSphere = v.znear / (cz - r);
    return depthSphere > depth;
}
#endif

}  // namespace bz
