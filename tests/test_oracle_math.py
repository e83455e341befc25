"""Pins the oracle's per-object arithmetic against independent formulations (CPU only).

Each check is written differently from the oracle so that a plausible mistake (a swapped operand, a dropped term,
a wrong sign or index, an off-by-one in the footprint) fails it:
  RotateQuat        float64 rotation matrix of the unit quaternion (ShaderBuffers.glsl:197-200)
  frustum           float64 signed distances to the four side planes and near/far (CullingShaderData.glsl:36-58)
  projectSphere     float64 sampling of the sphere's silhouette (Mara & McGuire 2013; CullingShaderData.glsl:8-33)
  LOD selection     float64 threshold, argmax over strictly smaller errors (CullingShaderData.glsl:103-116)
  Hi-Z level        numpy.frexp exponent (floor(log2) exactly), clamped like the sampler's LOD clamp
  Hi-Z footprint    texels with non-zero bilinear weight |u*W - (i + 0.5)| < 1, edge-clamped
  pyramid           numpy 2x2 block minimum for power-of-two levels; the bilinear-support rule for level 0
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import helpers as H

L = orc.load()


def fp(a):
    return np.ascontiguousarray(a, np.float32)


def rotate(v, q):
    v, q = fp(v), fp(q)
    out = np.zeros(3, np.float32)
    L.orc_rotate_quat(v.ctypes.data, q.ctypes.data, out.ctypes.data)
    return out


def quat_matrix64(q):
    x, y, z, w = (float(t) for t in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ------------------------------------------------------------------------------------------------- RotateQuat
def test_rotate_identity_exact():
    rng = np.random.default_rng(1)
    for _ in range(200):
        v = rng.uniform(-100, 100, 3).astype(np.float32)
        np.testing.assert_array_equal(rotate(v, [0, 0, 0, 1]), v)


def test_rotate_matches_rotation_matrix():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = rng.uniform(-2, 2, 3)
        got = rotate(v.astype(np.float32), q.astype(np.float32)).astype(np.float64)
        exp = quat_matrix64(q.astype(np.float32)) @ v.astype(np.float32).astype(np.float64)
        assert np.allclose(got, exp, rtol=0, atol=2e-5 * (1 + np.abs(exp).max())), (q, v, got, exp)


def test_rotate_quarter_turns():
    s = math.sqrt(0.5)
    np.testing.assert_allclose(rotate([1, 0, 0], [0, 0, s, s]), [0, 1, 0], atol=1e-6)   # +90 deg about z
    np.testing.assert_allclose(rotate([0, 1, 0], [s, 0, 0, s]), [0, 0, 1], atol=1e-6)   # +90 deg about x
    np.testing.assert_allclose(rotate([0, 0, 1], [0, s, 0, s]), [1, 0, 0], atol=1e-6)   # +90 deg about y


# ------------------------------------------------------------------------------------------------- frustum
def frustum(c, r, view):
    f = orc.view_floats(view)
    c = fp(c)
    return bool(L.orc_frustum(c.ctypes.data, r, f[35], f[36], f[37], f[38], f[41], f[42]))


@pytest.mark.parametrize("fov,w,h", [(70.0, 1920, 1080), (50.0, 512, 512), (90.0, 2048, 1024)])
def test_frustum_matches_plane_distances(fov, w, h):
    v = H.view_bytes(H.view(fov=fov, w=w, h=h))
    ty = math.tan(math.radians(fov) / 2)
    tx = ty * w / h
    # inward unit normals of the side planes through the eye (x = +-tx z, y = +-ty z)
    normals = [np.array([-1, 0, tx]) / math.hypot(1, tx), np.array([1, 0, tx]) / math.hypot(1, tx),
               np.array([0, -1, ty]) / math.hypot(1, ty), np.array([0, 1, ty]) / math.hypot(1, ty)]
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(20000):
        c = rng.uniform([-900, -900, -100], [900, 900, 900])
        r = rng.uniform(0.1, 40)
        d = [float(n @ c) for n in normals] + [c[2] - 0.1, 650.0 - c[2]]
        if min(abs(x + r) for x in d) < 1e-2 * (1 + abs(c).max()):
            continue  # within rounding of a plane
        exp = all(x > -r for x in d)
        assert frustum(c, r, v) == exp, (c, r, d)
        checked += 1
    assert checked > 15000


# ------------------------------------------------------------------------------------------------- projectSphere
def project(c, r, znear, p00, p11):
    c = fp(c)
    a = np.zeros(4, np.float32)
    ok = L.orc_project_sphere(c.ctypes.data, r, znear, p00, p11, a.ctypes.data)
    return bool(ok), a


def test_project_sphere_bounds_silhouette():
    v = H.view()
    p00, p11 = v.proj0, v.proj5
    rng = np.random.default_rng(4)
    for _ in range(400):
        c = rng.uniform([-50, -50, 5], [50, 50, 200])
        r = rng.uniform(0.2, 3.0)
        ok, a = project(c, r, 0.1, p00, p11)
        assert ok
        # dense sample of the sphere surface, projected with the infinite perspective into UV (v down)
        u = rng.normal(size=(4000, 3))
        pts = c + r * u / np.linalg.norm(u, axis=1, keepdims=True)
        U = pts[:, 0] / pts[:, 2] * p00 * 0.5 + 0.5
        V = -pts[:, 1] / pts[:, 2] * p11 * 0.5 + 0.5
        tol = 1e-5
        assert a[0] <= U.min() + tol and a[2] >= U.max() - tol
        assert a[1] <= V.min() + tol and a[3] >= V.max() - tol
        # tight: the extreme samples reach the box to within the sampling density
        span = max(a[2] - a[0], a[3] - a[1])
        assert U.min() - a[0] < 0.02 * span + 1e-4 and a[2] - U.max() < 0.02 * span + 1e-4
        assert V.min() - a[1] < 0.02 * span + 1e-4 and a[3] - V.max() < 0.02 * span + 1e-4


def test_project_sphere_near_plane_rule():
    ok, _ = project([0, 0, 1.0], 0.95, 0.1, 1.0, 1.0)   # c.z < r + znear
    assert not ok
    ok, _ = project([0, 0, 1.06], 0.95, 0.1, 1.0, 1.0)
    assert ok


# ------------------------------------------------------------------------------------------------- LOD selection
def test_lod_select_last_strictly_smaller_error():
    lods = np.zeros(16, orc_lod := np.dtype([("ic", "<u4"), ("fi", "<u4"), ("co", "<u4"), ("cc", "<u4"),
                                             ("err", "<f4"), ("pad", "<u4", 3)]))
    lods["err"][:8] = [0.01 * 2 ** k for k in range(8)]
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(5000):
        c = rng.uniform(-800, 800, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.3, 3))
        scale = np.float32(rng.uniform(0.5, 2))
        lt = np.float32(rng.uniform(1e-4, 5e-3))
        cnt = int(rng.integers(1, 9))
        dist = max(float(np.linalg.norm(c.astype(np.float64))) - float(r), 0.0)
        thr = dist * float(lt) / float(scale)
        errs = [float(np.float32(e)) for e in lods["err"][:cnt]]
        if any(abs(e - thr) < 1e-5 * max(thr, 1e-6) for e in errs):
            continue  # tie within rounding
        exp = max([i for i in range(1, cnt) if errs[i] < thr], default=0)
        got = L.orc_lod_select(c.ctypes.data, float(r), float(scale), float(lt), 0, cnt, lods.ctypes.data)
        assert got == exp, (c, r, scale, lt, cnt, thr)
        checked += 1
    assert checked > 4500


def test_lod_select_offset_and_single_lod():
    lods = np.zeros(12, np.dtype([("ic", "<u4"), ("fi", "<u4"), ("co", "<u4"), ("cc", "<u4"), ("err", "<f4"),
                                  ("pad", "<u4", 3)]))
    lods["err"][4:8] = [0.0, 0.0, 0.0, 0.0]
    c = fp([0, 0, 100])
    assert L.orc_lod_select(c.ctypes.data, 1.0, 1.0, 0.01, 4, 4, lods.ctypes.data) == 3  # all errors below: last wins
    assert L.orc_lod_select(c.ctypes.data, 1.0, 1.0, 0.01, 4, 1, lods.ctypes.data) == 0  # one LOD: index 0
    lods["err"][4:8] = [0.0, 9.0, 9.0, 9.0]
    assert L.orc_lod_select(c.ctypes.data, 1.0, 1.0, 0.01, 4, 4, lods.ctypes.data) == 0


# ------------------------------------------------------------------------------------------------- Hi-Z level
def test_hiz_level_is_floor_log2_clamped():
    rng = np.random.default_rng(6)
    vals = np.concatenate([rng.uniform(0, 4096, 20000), 2.0 ** rng.integers(-10, 20, 2000),
                           np.nextafter(2.0 ** rng.integers(1, 20, 2000), 0),
                           [0.0, -1.0, 0.5, 1.0, 1e-40, np.inf, np.nan, 3.4e38]]).astype(np.float32)
    for mips in (1, 5, 11, 16):
        for m in vals:
            if not (m >= 1.0):
                exp = 0
            else:
                _, e = np.frexp(np.float64(m)) if np.isfinite(m) else (0, 129)
                exp = min(int(e) - 1, mips - 1)
            assert L.orc_hiz_level(float(m), mips) == exp, (m, mips)


# ------------------------------------------------------------------------------------------------- footprint
def footprint_support(level, W, Hh, u, v):
    """Texels whose bilinear weight is non-zero: |t - (i + 0.5)| < 1 (edge-clamped), likewise for y, where
    t = u*W is the sampler's fp32 product (one rounding); the rest is exact arithmetic."""
    tx, ty = float(np.float32(u) * np.float32(W)), float(np.float32(v) * np.float32(Hh))
    xs = {min(max(i, 0), W - 1) for i in range(math.floor(tx) - 2, math.floor(tx) + 3) if abs(tx - (i + 0.5)) < 1}
    ys = {min(max(j, 0), Hh - 1) for j in range(math.floor(ty) - 2, math.floor(ty) + 3) if abs(ty - (j + 0.5)) < 1}
    return min(level[y, x] for y in ys for x in xs)


def test_footprint_min_is_bilinear_support():
    rng = np.random.default_rng(8)
    for W, Hh in [(1024, 512), (7, 3), (1, 1), (2, 1), (1, 16), (33, 17)]:
        level = rng.uniform(0, 1, (Hh, W)).astype(np.float32)
        lv = np.ascontiguousarray(level)
        samples = list(rng.uniform(-0.2, 1.2, (3000, 2)))
        # exact texel centres and edges (zero-weight neighbours), where u*W - 0.5 is an integer
        samples += [((i + 0.5) / W, (j + 0.5) / Hh) for i in range(min(W, 9)) for j in range(min(Hh, 9))]
        samples += [(i / W, j / Hh) for i in range(min(W, 9)) for j in range(min(Hh, 9))]
        for u, v in samples:
            u, v = np.float32(u), np.float32(v)
            got = L.orc_footprint_min(lv.ctypes.data, W, Hh, float(u), float(v))
            assert got == footprint_support(level, W, Hh, u, v), (W, Hh, u, v)


# ------------------------------------------------------------------------------------------------- pyramid
def test_pyramid_levels_are_block_minima():
    rng = np.random.default_rng(9)
    depth = rng.uniform(0, 1, (1080, 1920)).astype(np.float32)
    bw, bh, mips = 1024, 512, 11
    tex = orc.pyramid_build(depth, bw, bh, mips)
    offs, dims, _ = _levels(bw, bh, mips)
    prev = None
    for Lv, ((w, h), off) in enumerate(zip(dims, offs)):
        cur = tex[off:off + w * h].reshape(h, w)
        if Lv == 0:
            ys = [((y + 0.5) / h) for y in range(0, h, 37)]
            xs = [((x + 0.5) / w) for x in range(0, w, 29)]
            for v in ys:
                for u in xs:
                    x, y = int(u * w), int(v * h)
                    assert cur[y, x] == footprint_support(depth, 1920, 1080, np.float32(u), np.float32(v))
        else:
            ph, pw = prev.shape
            p2 = prev if pw > 1 else np.repeat(prev, 2, axis=1)
            p2 = p2 if ph > 1 else np.repeat(p2, 2, axis=0)
            exp = p2.reshape(h, 2, w, 2).min(axis=(1, 3))
            np.testing.assert_array_equal(cur, exp)
        prev = cur


def _levels(bw, bh, mips):
    offs, dims, off = [], [], 0
    for Lv in range(mips):
        w, h = max(1, bw >> Lv), max(1, bh >> Lv)
        offs.append(off)
        dims.append((w, h))
        off += w * h
    return offs, dims, off


# ------------------------------------------------------------------------------------------------- occlusion
def test_occlusion_against_constant_wall():
    """A constant-depth pyramid d0: an object is visible iff its nearest depth zNear / (z - r) exceeds d0."""
    bw, bh, mips = 256, 128, 9
    d0 = np.float32(0.01)  # wall at view depth zNear / d0 = 10
    tex = np.full(_levels(bw, bh, mips)[2], d0, np.float32)
    pyr = orc.pyramid(tex, bw, bh, mips)
    rng = np.random.default_rng(10)
    for _ in range(3000):
        c = np.array([rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(1, 40)], np.float32)
        r = np.float32(rng.uniform(0.05, 2))
        ok, a = project(c, r, 0.1, 1.5, 2.0)
        if not ok:
            continue
        near = 0.1 / (float(c[2]) - float(r))
        if abs(near - float(d0)) < 1e-6:
            continue
        got = L.orc_occlusion(a.ctypes.data, C.byref(pyr), float(bw), float(bh), c.ctypes.data, float(r), 0.1)
        assert bool(got) == (near > float(d0))


def test_occlusion_selects_level_by_footprint_size():
    """Each level holds its own constant d_L: the answer reveals which level was read. The expected level is
    floor(log2(max(w, h))) in float64 (clamped), w, h the UV box times the pyramid's extent."""
    bw, bh, mips = 1024, 512, 11
    offs, dims, total = _levels(bw, bh, mips)
    d = np.array([0.002 * (Lv + 1) for Lv in range(mips)], np.float32)
    tex = np.zeros(total, np.float32)
    for Lv, off in enumerate(offs):
        tex[off:off + dims[Lv][0] * dims[Lv][1]] = d[Lv]
    pyr = orc.pyramid(tex, bw, bh, mips)
    rng = np.random.default_rng(11)
    hit = set()
    for _ in range(6000):
        c = np.array([rng.uniform(-5, 5), rng.uniform(-3, 3), rng.uniform(0.5, 60)], np.float32)
        r = np.float32(np.exp(rng.uniform(np.log(0.002), np.log(3))))
        ok, a = project(c, r, 0.1, 1.4281, 2.5389)
        if not ok:
            continue
        m = max((float(a[2]) - float(a[0])) * bw, (float(a[3]) - float(a[1])) * bh)
        if m >= 1 and abs(math.log2(m) - round(math.log2(m))) < 1e-4:
            continue
        lvl = 0 if m < 1 else min(int(math.floor(math.log2(m))), mips - 1)
        near = 0.1 / (float(c[2]) - float(r))
        if abs(near - float(d[lvl])) < 1e-6 * near:
            continue
        got = L.orc_occlusion(a.ctypes.data, C.byref(pyr), float(bw), float(bh), c.ctypes.data, float(r), 0.1)
        assert bool(got) == (near > float(d[lvl])), (c, r, m, lvl)
        hit.add(lvl)
    assert len(hit) >= 8


def test_occlusion_samples_box_centre():
    """Level 0 split in half (0 on one side, 1 on the other), along u and then along v; a small box decides by
    which half its centre lies in (v runs down the screen)."""
    bw, bh, mips = 1024, 1024, 11
    offs, dims, total = _levels(bw, bh, mips)
    for axis in (0, 1):
        tex = np.zeros(total, np.float32)
        lvl0 = tex[:bw * bh].reshape(bh, bw)
        if axis == 0:
            lvl0[:, bw // 2:] = 1.0
        else:
            lvl0[bh // 2:, :] = 1.0
        pyr = orc.pyramid(tex, bw, bh, mips)
        checked = 0
        for x in np.linspace(-4.0, 4.0, 401):
            c = np.array([x, 0.0, 10.0] if axis == 0 else [0.0, x, 10.0], np.float32)
            ok, a = project(c, np.float32(0.001), 0.1, 1.0, 1.0)
            cu = (float(a[0]) + float(a[2])) / 2 if axis == 0 else (float(a[1]) + float(a[3])) / 2
            if abs(cu * bw - bw / 2) < 1.0:
                continue  # footprint straddles the split
            got = L.orc_occlusion(a.ctypes.data, C.byref(pyr), float(bw), float(bh), c.ctypes.data, 0.001, 0.1)
            assert bool(got) == (cu < 0.5), (axis, x, cu)
            checked += 1
        assert checked > 390


# ------------------------------------------------------------------------------------------------- rounding at planes
def unfused_frustum(objs, xfs, surfs, v):
    """The transform and frustum test in numpy float32 with every product and sum rounded on its own (no fused
    multiply-add anywhere): the arithmetic the cull would have without the fused points."""
    f = np.float32
    s = surfs[objs["surfaceId"]]
    vx, vy, vz = (s["center"][:, k].astype(f) for k in range(3))
    x, y, z, w = (xfs["orientation"][:, k].astype(f) for k in range(4))
    tx = (y * vz - z * vy) + w * vx
    ty = (z * vx - x * vz) + w * vy
    tz = (x * vy - y * vx) + w * vz
    rx, ry, rz = vx + f(2) * (y * tz - z * ty), vy + f(2) * (z * tx - x * tz), vz + f(2) * (x * ty - y * tx)
    sig = xfs["scale"].astype(f)
    wx, wy, wz = (rr * sig + xfs["pos"][:, k].astype(f) for k, rr in enumerate((rx, ry, rz)))
    m = np.array(v.view, f)
    cx, cy, cz = (((m[k] * wx + wy * m[4 + k]) + wz * m[8 + k]) + f(1) * m[12 + k] for k in range(3))
    r = s["radius"].astype(f) * sig
    fR, fL, fT, fB = (f(t) for t in (v.frustumRight, v.frustumLeft, v.frustumTop, v.frustumBottom))
    return ((cz * fL - np.abs(cx) * fR > -r) & (cz * fB - np.abs(cy) * fT > -r) & (cz + r > f(v.zNear))
            & (cz - r < f(v.zFar)))


def unfused_hiz_level(objs, xfs, surfs, v, mips=11):
    """The Hi-Z level of each object's projected box in numpy float32 without fused multiply-adds (transform,
    projectSphere, box side in texels, floor(log2) by frexp)."""
    f = np.float32
    s = surfs[objs["surfaceId"]]
    vx, vy, vz = (s["center"][:, k].astype(f) for k in range(3))
    x, y, z, w = (xfs["orientation"][:, k].astype(f) for k in range(4))
    tx = (y * vz - z * vy) + w * vx
    ty = (z * vx - x * vz) + w * vy
    tz = (x * vy - y * vx) + w * vz
    rx, ry, rz = vx + f(2) * (y * tz - z * ty), vy + f(2) * (z * tx - x * tz), vz + f(2) * (x * ty - y * tx)
    sig = xfs["scale"].astype(f)
    wx, wy, wz = (rr * sig + xfs["pos"][:, k].astype(f) for k, rr in enumerate((rx, ry, rz)))
    m = np.array(v.view, f)
    cx, cy, cz = (((m[k] * wx + wy * m[4 + k]) + wz * m[8 + k]) + f(1) * m[12 + k] for k in range(3))
    r = s["radius"].astype(f) * sig
    crx, cry, crz = cx * r, cy * r, cz * r
    czr2 = cz * cz - r * r
    sx, sy = np.sqrt(cx * cx + czr2), np.sqrt(cy * cy + czr2)
    minx, maxx = (sx * cx - crz) / (sx * cz + crx), (sx * cx + crz) / (sx * cz - crx)
    miny, maxy = (sy * cy - crz) / (sy * cz + cry), (sy * cy + crz) / (sy * cz - cry)
    p00, p11 = f(v.proj0), f(v.proj5)
    a0, a2 = (minx * p00) * f(0.5) + f(0.5), (maxx * p00) * f(0.5) + f(0.5)
    a1, a3 = (maxy * p11) * f(-0.5) + f(0.5), (miny * p11) * f(-0.5) + f(0.5)
    side = np.maximum((a2 - a0) * f(v.pyramidWidth), (a3 - a1) * f(v.pyramidHeight))
    e = np.frexp(side.astype(np.float64))[1] - 1
    return np.where(side >= 1, np.minimum(e, mips - 1), 0)


def test_hiz_level_scene_separates_the_fused_points():
    """tests.helpers.hiz_level_scene puts each projected box's side just below a power of two in texels. Against the
    pyramid of tests.helpers.parity_pyramid (0.0 on even levels, 1.0 on odd ones) an object passes the Hi-Z test
    exactly when its level is even, so the TEMPORAL pass reveals every level choice: the oracle's (fused points) and
    the unfused arithmetic's differ on about 1% of the objects, and both levels occur often. The GPU parity tests on
    this scene (tests/test_gpu_boundary.py) therefore fail for a projection that fuses at other points."""
    tex = H.parity_pyramid(1024, 512, 11)
    opyr = orc.pyramid(tex, 1024, 512, 11)
    for yaw, pitch, seed in ((0.3, 0.1, 7), (-1.1, -0.4, 8)):
        v = H.view(yaw=yaw, pitch=pitch, pyr=(1024, 512))
        objs, xfs, surfs, lods = H.hiz_level_scene(20000, v, seed)
        cmds, cnt = orc.cull((objs, xfs, surfs, lods), H.view_bytes(v), 3, opyr, capacity=20000)  # TEMPORAL
        fused = np.zeros(20000, bool)
        fused[H.cmds_u32(cmds)[:cnt, 0]] = True
        assert 0.1 * 20000 < cnt < 0.9 * 20000, cnt
        plain = (unfused_hiz_level(objs, xfs, surfs, v) % 2 == 0) & unfused_frustum(objs, xfs, surfs, v)
        diff = int(np.count_nonzero(plain != fused))
        assert 100 < diff < 2000, diff


def test_boundary_scene_separates_the_fused_points():
    """The objects of tests.helpers.boundary_scene sit on the frustum's side planes, so their decisions depend on
    where the arithmetic rounds: about half are inside, and the oracle (fused multiply-adds at its fixed points) and
    the same arithmetic without fusion decide a few percent of them differently. A kernel that fused at other points,
    or not at all, therefore fails the GPU parity test on this scene (tests/test_gpu_boundary.py)."""
    for yaw, pitch, seed in ((0.3, 0.1, 7), (-1.1, -0.4, 8)):
        v = H.view(yaw=yaw, pitch=pitch)
        objs, xfs, surfs, lods = H.boundary_scene(20000, v, seed)
        cmds, cnt = orc.cull((objs, xfs, surfs, lods), H.view_bytes(v), 2, capacity=20000)  # PLAIN
        fused = np.zeros(20000, bool)
        fused[H.cmds_u32(cmds)[:cnt, 0]] = True
        assert 0.25 * 20000 < cnt < 0.85 * 20000, cnt
        plain = unfused_frustum(objs, xfs, surfs, v)
        diff = int(np.count_nonzero(plain != fused))
        assert 200 < diff < 4000, diff
