"""Shared workload builders for the tests (seeded, identical for the oracle and the GPU path)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from blitzen_amd import cull as bz

SEED_C0 = 0xB1172E00
SEED_C1 = 0xB1172E01
SEED_C2 = 0xB1172E02
SEED_C4 = 0xB1172E04


def as_tuple(sc: bz.SynthScene):
    return (sc.objects, sc.transforms, sc.surfaces, sc.lods)


def c0_scene():
    """16,384 objects, 64 surfaces x 4 LODs, positions U[-500,500]^3, scale U[0.5,2]."""
    return bz.synth_scene(SEED_C0, 16384, 64, 4)


def c1_scene(n: int = 4194304, seed: int = SEED_C1):
    """Config-1 distributions: n objects over 256 surfaces x 8 LODs."""
    return bz.synth_scene(seed, n, 256, 8)


def view(fov=70.0, w=1920, h=1080, znear=0.1, zfar=650.0, pos=(0.0, 0.0, 0.0), yaw=0.0, pitch=0.0, pyr=(0, 0)):
    v = bz.camera_setup(fov, w, h, znear, zfar, pos, yaw, pitch)
    v.pyramidWidth = float(pyr[0])
    v.pyramidHeight = float(pyr[1])
    return v


def view_bytes(v) -> bytes:
    return bytes(memoryview(v))


def c1_depth(seed: int = SEED_C1, w=1920, h=1080, occluders=512, p11=None, znear=0.1):
    if p11 is None:
        p11 = view(w=w, h=h).proj5
    return bz.synth_depth(seed, w, h, occluders, p11, znear)


def cmds_u32(structured: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(structured).view(np.uint32).reshape(-1, 6)


def boundary_scene(n: int, v, seed: int):
    """n objects (16 surfaces x 4 LODs) whose view-space bounding spheres touch a side plane of `v`'s frustum in
    exact arithmetic: even objects a left/right plane, odd ones a top/bottom plane, depth U[20, 300], each placed
    from float64 arithmetic and rounded to float32 positions, so that whether an object is inside depends on the
    rounding of every step of the transform and the plane test."""
    sc = bz.synth_scene(seed, n, 16, 4)
    objs, xfs, surfs, lods = sc.objects, sc.transforms.copy(), sc.surfaces, sc.lods
    rng = np.random.default_rng(seed)
    s = surfs[objs["surfaceId"]]
    q = xfs["orientation"].astype(np.float64)
    sig = xfs["scale"].astype(np.float64)
    r = (s["radius"].astype(np.float32) * xfs["scale"].astype(np.float32)).astype(np.float64)
    c0 = s["center"].astype(np.float64)
    qv, qw = q[:, :3], q[:, 3:4]
    o = (c0 + 2.0 * np.cross(qv, np.cross(qv, c0) + qw * c0)) * sig[:, None]  # rotated, scaled centre
    m = np.array(v.view, np.float64).reshape(4, 4)  # m[col][row]
    V3, t = m[:3, :3].T, m[3, :3]
    fR, fL, fT, fB = (float(x) for x in (v.frustumRight, v.frustumLeft, v.frustumTop, v.frustumBottom))
    cz = rng.uniform(20.0, 300.0, n)
    sgn = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    inner = rng.uniform(-0.3, 0.3, n)
    xplane = np.arange(n) % 2 == 0
    cx = np.where(xplane, sgn * (cz * fL + r) / fR, inner * cz * fL / fR)
    cy = np.where(xplane, inner * cz * fB / fT, sgn * (cz * fB + r) / fT)
    world = (np.stack([cx, cy, cz], 1) - t) @ V3  # V3^T (c - t): V3 is a rotation
    xfs["pos"] = (world - o).astype(np.float32)
    return objs, xfs, surfs, lods


def hiz_level_scene(n: int, v, seed: int):
    """n objects whose view-space spheres sit on the view axis (x = y = 0 in exact arithmetic, depth U[20, 200]) with
    radii such that the projected box's larger side, in pyramid texels of `v` (pyramidWidth x pyramidHeight set), is
    within two rounding quanta below 2^k for odd k in {1, 3, 5, 7}: the Hi-Z level (floor of its log2) is k or k - 1
    depending on how each step of the transform and the projection rounds. Scales are chosen per object to give the
    radius."""
    sc = bz.synth_scene(seed, n, 16, 4)
    objs, xfs, surfs, lods = sc.objects, sc.transforms.copy(), sc.surfaces, sc.lods
    rng = np.random.default_rng(seed)
    s = surfs[objs["surfaceId"]]
    big = max(float(v.proj0) * float(v.pyramidWidth), float(v.proj5) * float(v.pyramidHeight))
    k = rng.choice(np.array([1, 3, 5, 7]), n)
    cz = rng.uniform(20.0, 200.0, n)
    # the side's fp32 value is quantised by the subtraction of two UV coordinates near 0.5 (2^-24 in UV, 2^-14 of a
    # texel at 1024 texels): aim up to two quanta below 2^k so that it rounds to either side
    side = np.exp2(k) - rng.uniform(0.0, 2.0, n) * 2.0 ** -14
    ratio = side / big  # r / sqrt(cz^2 - r^2) for a side of `side` texels
    r_target = cz * ratio / np.sqrt(1.0 + ratio * ratio)
    xfs["scale"] = (r_target / s["radius"].astype(np.float64)).astype(np.float32)
    q = xfs["orientation"].astype(np.float64)
    sig = xfs["scale"].astype(np.float64)
    c0 = s["center"].astype(np.float64)
    qv, qw = q[:, :3], q[:, 3:4]
    o = (c0 + 2.0 * np.cross(qv, np.cross(qv, c0) + qw * c0)) * sig[:, None]
    m = np.array(v.view, np.float64).reshape(4, 4)
    V3, t = m[:3, :3].T, m[3, :3]
    world = (np.stack([np.zeros(n), np.zeros(n), cz], 1) - t) @ V3
    xfs["pos"] = (world - o).astype(np.float32)
    return objs, xfs, surfs, lods


def parity_pyramid(base_w: int, base_h: int, mips: int) -> np.ndarray:
    """Pyramid texels (levels back to back) holding 0.0 on even levels and 1.0 on odd ones: a sphere passes the Hi-Z
    test exactly when its level is even."""
    out = []
    for L in range(mips):
        w, h = max(1, base_w >> L), max(1, base_h >> L)
        out.append(np.full(w * h, float(L & 1), np.float32))
    return np.concatenate(out + [np.zeros(1, np.float32)])
