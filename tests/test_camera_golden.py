"""Camera constants and BlitML operations: oracle and product vs golden vectors from the reference's own code.

The fixture tests/golden/camera_golden.json was written by tests/golden/make_golden.py from the reference's
SetupCamera/UpdateProjection (src/Game/blitzenCamera.cpp:14-41, 93-126) and BlitML (src/BlitzenMathLibrary/blitML.h)
compiled from their sources into oracle/_ref. Comparisons are bit-exact.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from blitzen_amd import cull as bz
from oracle import pyoracle as orc

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "camera_golden.json").read_text())
MASK = np.ones(46, bool)
MASK[[43, 44]] = False  # pyramidWidth/Height are not set by SetupCamera


def u32(a):
    return np.asarray(a, np.uint32)


def f32(bits_list):
    return u32(bits_list).view(np.float32)


@pytest.mark.parametrize("case", GOLD["cameras"], ids=lambda c: f"fov{c['fov_deg']:.1f}_{int(c['w'])}x{int(c['h'])}")
def test_oracle_camera_matches_reference(case):
    fov = float(f32([case["fov_rad_bits"]])[0])
    out = orc.camera_rad(fov, case["w"], case["h"], case["znear"], case["zfar"], tuple(case["pos"]), case["yaw"],
                         case["pitch"])
    got = np.frombuffer(out, np.uint32)[:46][MASK]
    np.testing.assert_array_equal(got, u32(case["view_data_bits"])[MASK])


@pytest.mark.parametrize("case", GOLD["cameras"], ids=lambda c: f"fov{c['fov_deg']:.1f}_{int(c['w'])}x{int(c['h'])}")
def test_product_camera_matches_reference(case):
    fov = float(f32([case["fov_rad_bits"]])[0])
    v = bz.CameraViewData()
    pos = (C.c_float * 3)(*case["pos"])
    assert bz.lib().bz_camera_setup(fov, case["w"], case["h"], case["znear"], case["zfar"], pos, case["yaw"],
                                    case["pitch"], C.byref(v)) == 0
    got = np.frombuffer(bytes(memoryview(v)), np.uint32)[:46][MASK]
    np.testing.assert_array_equal(got, u32(case["view_data_bits"])[MASK])


def test_radians_matches_reference():
    for case in GOLD["cameras"]:
        assert np.float32(orc.load().orc_radians(case["fov_deg"])).view(np.uint32) == case["fov_rad_bits"]


def test_blitml_ops_match_reference():
    lib = orc.load()
    for c in GOLD["ops"]["mat4_inverse"]:
        m = f32(c["in"]).copy()
        r = np.zeros(16, np.float32)
        lib.orc_mat4_inverse(m.ctypes.data, r.ctypes.data)
        np.testing.assert_array_equal(r.view(np.uint32), u32(c["out"]))
    for c in GOLD["ops"]["quat_to_mat4"]:
        q = f32(c["in"]).copy()
        r = np.zeros(16, np.float32)
        lib.orc_quat_to_mat4(q.ctypes.data, r.ctypes.data)
        np.testing.assert_array_equal(r.view(np.uint32), u32(c["out"]))
    for c in GOLD["ops"]["quat_from_angle_axis"]:
        ax = f32(c["axis"]).copy()
        q = np.zeros(4, np.float32)
        lib.orc_quat_from_angle_axis(ax.ctypes.data, float(f32([c["angle"]])[0]), q.ctypes.data)
        np.testing.assert_array_equal(q.view(np.uint32), u32(c["out"]))


def test_mat_vec_order_matches_reference():
    """The view transform's sum order (m0*x + y*m4 + z*m8 + w*m12, blitMLTypes.h:184-193) against the reference's
    mat4 * vec4, through the oracle's sphere transform with an identity rotation and unit scale."""
    lib = orc.load()
    for c in GOLD["ops"]["mat4_mul_vec4"]:
        m, v = f32(c["m"]).copy(), f32(c["v"]).copy()
        if v[3] != 0:
            v3 = v[:3].copy()
            # w == 1 for the cull's transform: compare against the reference with w = 1
            ref = orc.load_ref()
            if ref is None:
                pytest.skip("oracle/_ref not built")
            vv = np.array([v3[0], v3[1], v3[2], 1.0], np.float32)
            exp = np.zeros(4, np.float32)
            ref.ref_mat4_mul_vec4(m.ctypes.data, vv.ctypes.data, exp.ctypes.data)
            out_c = np.zeros(3, np.float32)
            out_r = C.c_float()
            q = np.array([0, 0, 0, 1], np.float32)
            lib.orc_sphere_to_view(v3.ctypes.data, 1.0, np.zeros(3, np.float32).ctypes.data, 1.0, q.ctypes.data,
                                   m.ctypes.data, out_c.ctypes.data, C.byref(out_r))
            np.testing.assert_array_equal(out_c.view(np.uint32), exp[:3].view(np.uint32))


def test_pyramid_extent_matches_reference():
    for c in GOLD["pyramid_extent"]:
        bw, bh, mips = bz.reference_pyramid_extent(c["w"], c["h"])
        assert [bw, bh] == c["prev_pow2"] and mips == c["mips"]
        assert orc.load().orc_previous_pow2(c["w"]) == c["prev_pow2"][0]
        assert orc.load().orc_mip_levels(bw, bh) == c["mips"]


def test_live_reference_agrees_on_random_poses():
    """Where oracle/_ref is built (the container with the reference tree), compare on fresh random poses too."""
    ref = orc.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(7)
    out = np.zeros(256, np.uint8)
    for _ in range(300):
        fov = float(np.float32(rng.uniform(0.3, 2.5)))
        w, h = float(rng.integers(16, 8192)), float(rng.integers(16, 8192))
        pos = rng.uniform(-5000, 5000, 3).astype(np.float32)
        yaw, pitch = float(np.float32(rng.uniform(-10, 10))), float(np.float32(rng.uniform(-3, 3)))
        ref.ref_setup_camera(fov, w, h, 0.1, 650.0, pos.ctypes.data, yaw, pitch, out.ctypes.data)
        exp = np.frombuffer(out.tobytes(), np.uint32)[:46][MASK]
        got = np.frombuffer(orc.camera_rad(fov, w, h, 0.1, 650.0, tuple(pos), yaw, pitch), np.uint32)[:46][MASK]
        np.testing.assert_array_equal(got, exp)
