// ref_camera.cpp — C entry points over the reference's own camera and BlitML code, compiled from the sources under
// the reference tree (src/Game/blitzenCamera.cpp, src/BlitzenMathLibrary/blitML.h) into oracle/_ref/. TEST
// INFRASTRUCTURE ONLY: used to write and check golden vectors for the camera constants and the BlitML operations
// the cull path restates. Nothing here is reimplemented; every function forwards to the reference.
#include <cstring>
#include "Game/blitCamera.h"

extern "C" {

// SetupCamera (blitzenCamera.cpp:14-41) -> the 256-byte CameraViewData (blitCamera.h:38-64).
void ref_setup_camera(float fov, float w, float h, float znear, float drawDistance, const float pos[3], float yaw,
                      float pitch, void* out256)
{
    BlitzenEngine::Camera cam;
    BlitzenEngine::SetupCamera(cam, fov, w, h, znear, BlitML::vec3(pos[0], pos[1], pos[2]), drawDistance, yaw, pitch);
    std::memcpy(out256, &cam.viewData, sizeof(cam.viewData));
}

float ref_radians(float deg) { return BlitML::Radians(deg); }

void ref_mat4_inverse(const float* m, float* out)
{
    BlitML::mat4 a;
    std::memcpy(a.data, m, 64);
    BlitML::mat4 r = BlitML::Mat4Inverse(a);
    std::memcpy(out, r.data, 64);
}

void ref_mat4_mul(const float* a, const float* b, float* out)
{
    BlitML::mat4 x, y;
    std::memcpy(x.data, a, 64);
    std::memcpy(y.data, b, 64);
    BlitML::mat4 r = x * y;
    std::memcpy(out, r.data, 64);
}

void ref_mat4_mul_vec4(const float* m, const float* v, float* out)
{
    BlitML::mat4 x;
    std::memcpy(x.data, m, 64);
    BlitML::vec4 r = x * BlitML::vec4(v[0], v[1], v[2], v[3]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void ref_quat_to_mat4(const float* q, float* out)
{
    BlitML::mat4 r = BlitML::QuatToMat4(BlitML::quat(q[0], q[1], q[2], q[3]));
    std::memcpy(out, r.data, 64);
}

void ref_quat_from_angle_axis(const float* axis, float angle, float* out)
{
    BlitML::quat q = BlitML::QuatFromAngleAxis(BlitML::vec3(axis[0], axis[1], axis[2]), angle, 0);
    out[0] = q.x; out[1] = q.y; out[2] = q.z; out[3] = q.w;
}

void ref_cross(const float* a, const float* b, float* out)
{
    BlitML::vec3 r = BlitML::Cross(BlitML::vec3(a[0], a[1], a[2]), BlitML::vec3(b[0], b[1], b[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void ref_normalize_plane(const float* p, float* out)
{
    BlitML::vec4 r = BlitML::NormalizePlane(BlitML::vec4(p[0], p[1], p[2], p[3]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

float ref_length3(const float* v) { return BlitML::Length(BlitML::vec3(v[0], v[1], v[2])); }

uint32_t ref_previous_pow2(uint32_t v) { return BlitML::PreviousPow2(v); }

uint32_t ref_mip_levels(uint32_t w, uint32_t h) { return BlitML::GetDepthPyramidMipLevels(w, h); }

uint32_t ref_group_size(uint32_t n, uint32_t per) { return BlitML::GetComputeShaderGroupSize(n, per); }

}
