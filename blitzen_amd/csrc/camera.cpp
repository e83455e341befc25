// camera.cpp — per-view cull constants (CameraViewData) from a camera pose.
//
// Same result, bit for bit, as Blitzen's SetupCamera + UpdateProjection (src/Game/blitzenCamera.cpp:14-41, 93-126):
// InfiniteZPerspective (blitML.h:261-278), view = Mat4Inverse(Translate(pos) * QuatToMat4(yaw) * QuatToMat4(pitch))
// (blitML.h:312-372, 520-534, 559-570), frustum planes from the normalised rows of the projection
// (blitML.h:223, blitMLTypes.h:143-170) and lodTarget = (2 / P11) * (1 / height). Every expression keeps BlitML's
// fp32 evaluation order; the file is compiled with -ffp-contract=off.
#include <cmath>
#include <cstring>

#include "../../include/bzcull.h"

namespace {

struct Mat4 {
    float d[16];
};

Mat4 identity()
{
    Mat4 m;
    std::memset(&m, 0, sizeof m);
    m.d[0] = m.d[5] = m.d[10] = m.d[15] = 1.0f;
    return m;
}

// res[j + 4i] = a[j] b[4i] + a[4+j] b[4i+1] + a[8+j] b[4i+2] + a[12+j] b[4i+3]  (column-major product)
Mat4 mul(const Mat4& a, const Mat4& b)
{
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = a.d[j] * b.d[4 * i];
            s = s + a.d[4 + j] * b.d[4 * i + 1];
            s = s + a.d[8 + j] * b.d[4 * i + 2];
            s = s + a.d[12 + j] * b.d[4 * i + 3];
            r.d[j + 4 * i] = s;
        }
    return r;
}

Mat4 quat_to_mat(const float q[4])
{
    float n2 = q[0] * q[0];
    n2 = n2 + q[1] * q[1];
    n2 = n2 + q[2] * q[2];
    n2 = n2 + q[3] * q[3];
    const float len = std::sqrt(n2);
    const float x = q[0] / len, y = q[1] / len, z = q[2] / len, w = q[3] / len;
    Mat4 r = identity();
    r.d[0] = (1.0f - 2.0f * y * y) - 2.0f * z * z;
    r.d[1] = 2.0f * x * y - 2.0f * z * w;
    r.d[2] = 2.0f * x * z + 2.0f * y * w;
    r.d[4] = 2.0f * x * y + 2.0f * z * w;
    r.d[5] = (1.0f - 2.0f * x * x) - 2.0f * z * z;
    r.d[6] = 2.0f * y * z - 2.0f * x * w;
    r.d[8] = 2.0f * x * z - 2.0f * y * w;
    r.d[9] = 2.0f * y * z + 2.0f * x * w;
    r.d[10] = (1.0f - 2.0f * x * x) - 2.0f * y * y;
    return r;
}

void angle_axis(float ax, float ay, float az, float angle, float q[4])
{
    const float h = 0.5f * angle;
    const float s = std::sin(h), c = std::cos(h);
    q[0] = s * ax;
    q[1] = s * ay;
    q[2] = s * az;
    q[3] = c;
}

float sum3(float a, float b, float c) { return (a + b) + c; }

// General 4x4 inverse by cofactors, in the term order of blitML.h:312-363.
Mat4 inverse(const Mat4& M)
{
    const float* m = M.d;
    const float t0 = m[10] * m[15], t1 = m[14] * m[11], t2 = m[6] * m[15], t3 = m[14] * m[7];
    const float t4 = m[6] * m[11], t5 = m[10] * m[7], t6 = m[2] * m[15], t7 = m[14] * m[3];
    const float t8 = m[2] * m[11], t9 = m[10] * m[3], t10 = m[2] * m[7], t11 = m[6] * m[3];
    const float t12 = m[8] * m[13], t13 = m[12] * m[9], t14 = m[4] * m[13], t15 = m[12] * m[5];
    const float t16 = m[4] * m[9], t17 = m[8] * m[5], t18 = m[0] * m[13], t19 = m[12] * m[1];
    const float t20 = m[0] * m[9], t21 = m[8] * m[1], t22 = m[0] * m[5], t23 = m[4] * m[1];
    Mat4 R;
    float* r = R.d;
    r[0] = sum3(t0 * m[5], t3 * m[9], t4 * m[13]) - sum3(t1 * m[5], t2 * m[9], t5 * m[13]);
    r[1] = sum3(t1 * m[1], t6 * m[9], t9 * m[13]) - sum3(t0 * m[1], t7 * m[9], t8 * m[13]);
    r[2] = sum3(t2 * m[1], t7 * m[5], t10 * m[13]) - sum3(t3 * m[1], t6 * m[5], t11 * m[13]);
    r[3] = sum3(t5 * m[1], t8 * m[5], t11 * m[9]) - sum3(t4 * m[1], t9 * m[5], t10 * m[9]);
    float det = m[0] * r[0];
    det = det + m[4] * r[1];
    det = det + m[8] * r[2];
    det = det + m[12] * r[3];
    const float d = 1.0f / det;
    for (int k = 0; k < 4; ++k) r[k] = d * r[k];
    r[4] = d * (sum3(t1 * m[4], t2 * m[8], t5 * m[12]) - sum3(t0 * m[4], t3 * m[8], t4 * m[12]));
    r[5] = d * (sum3(t0 * m[0], t7 * m[8], t8 * m[12]) - sum3(t1 * m[0], t6 * m[8], t9 * m[12]));
    r[6] = d * (sum3(t3 * m[0], t6 * m[4], t11 * m[12]) - sum3(t2 * m[0], t7 * m[4], t10 * m[12]));
    r[7] = d * (sum3(t4 * m[0], t9 * m[4], t10 * m[8]) - sum3(t5 * m[0], t8 * m[4], t11 * m[8]));
    r[8] = d * (sum3(t12 * m[7], t15 * m[11], t16 * m[15]) - sum3(t13 * m[7], t14 * m[11], t17 * m[15]));
    r[9] = d * (sum3(t13 * m[3], t18 * m[11], t21 * m[15]) - sum3(t12 * m[3], t19 * m[11], t20 * m[15]));
    r[10] = d * (sum3(t14 * m[3], t19 * m[7], t22 * m[15]) - sum3(t15 * m[3], t18 * m[7], t23 * m[15]));
    r[11] = d * (sum3(t17 * m[3], t20 * m[7], t23 * m[11]) - sum3(t16 * m[3], t21 * m[7], t22 * m[11]));
    r[12] = d * (sum3(t14 * m[10], t17 * m[14], t13 * m[6]) - sum3(t16 * m[14], t12 * m[6], t15 * m[10]));
    r[13] = d * (sum3(t20 * m[14], t12 * m[2], t19 * m[10]) - sum3(t18 * m[10], t21 * m[14], t13 * m[2]));
    r[14] = d * (sum3(t18 * m[6], t23 * m[14], t15 * m[2]) - sum3(t22 * m[14], t14 * m[2], t19 * m[6]));
    r[15] = d * (sum3(t22 * m[10], t16 * m[2], t21 * m[6]) - sum3(t20 * m[6], t23 * m[10], t17 * m[2]));
    return R;
}

void normalize_plane(const float p[4], float out[4])
{
    const float len = std::sqrt(sum3(p[0] * p[0], p[1] * p[1], p[2] * p[2]));
    for (int k = 0; k < 4; ++k) out[k] = p[k] / len;
}

}  // namespace

extern "C" bz_status bz_camera_setup(float fov_radians, float window_w, float window_h, float z_near,
                                     float draw_distance, const float position[3], float yaw, float pitch,
                                     bz_CameraViewData* out)
{
    if (!out || !position || !(window_w > 0.0f) || !(window_h > 0.0f)) return BZ_E_INVALID;
    std::memset(out, 0, sizeof *out);
    out->zNear = z_near;
    out->zFar = draw_distance;
    for (int k = 0; k < 3; ++k) out->position[k] = position[k];
    Mat4 T = identity();
    T.d[12] = position[0];
    T.d[13] = position[1];
    T.d[14] = position[2];
    float qy[4], qp[4];
    angle_axis(0.0f, -1.0f, 0.0f, yaw, qy);
    angle_axis(1.0f, 0.0f, 0.0f, pitch, qp);
    const Mat4 R = mul(quat_to_mat(qy), quat_to_mat(qp));
    const Mat4 V = inverse(mul(T, R));
    std::memcpy(out->view, V.d, sizeof V.d);
    Mat4 P;
    std::memset(&P, 0, sizeof P);
    const float f = 1.0f / std::tan(fov_radians / 2.0f);
    const float aspect = window_w / window_h;
    P.d[0] = f / aspect;
    P.d[5] = f;
    P.d[11] = 1.0f;
    P.d[14] = z_near;
    out->proj0 = P.d[0];
    out->proj5 = P.d[5];
    const Mat4 PV = mul(P, V);
    std::memcpy(out->projectionView, PV.d, sizeof PV.d);
    const float px[4] = { P.d[3] + P.d[0], P.d[7] + P.d[4], P.d[11] + P.d[8], P.d[15] + P.d[12] };
    const float py[4] = { P.d[3] + P.d[1], P.d[7] + P.d[5], P.d[11] + P.d[9], P.d[15] + P.d[13] };
    float fx[4], fy[4];
    normalize_plane(px, fx);
    normalize_plane(py, fy);
    out->frustumRight = fx[0];
    out->frustumLeft = fx[2];
    out->frustumTop = fy[1];
    out->frustumBottom = fy[2];
    out->lodTarget = (2.0f / out->proj5) * (1.0f / window_h);
    return BZ_OK;
}
