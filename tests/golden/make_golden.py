"""Writes tests/golden/camera_golden.json from the reference's own compiled code (oracle/_ref/libref_camera.so).

Run in a container where the reference tree is present and `make -C oracle ref` has built oracle/_ref:
    python tests/golden/make_golden.py
The fixture holds inputs and outputs only (floats as IEEE-754 bit patterns), no reference source.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import pyoracle as orc  # noqa: E402

OUT = Path(__file__).resolve().parent / "camera_golden.json"


def bits(a) -> list[int]:
    return [int(x) for x in np.asarray(a, np.float32).view(np.uint32).ravel()]


def main():
    ref = orc.load_ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_camera.so is not built (needs the reference tree)")
    rng = np.random.default_rng(20261202)
    cases = []
    # the configured cameras: 70 deg at the origin looking +z (1920x1080), the reference's defaults
    # (70 deg, 1280x720, position (20, 70, 0), zFar 650), and seeded poses / fields of view
    fixed = [(70.0, 1920.0, 1080.0, 0.1, 650.0, (0.0, 0.0, 0.0), 0.0, 0.0),
             (70.0, 1280.0, 720.0, 0.1, 650.0, (20.0, 70.0, 0.0), 0.0, 0.0),
             (50.0, 512.0, 512.0, 0.1, 1000.0, (0.0, 0.0, 0.0), 0.0, 0.0),
             (90.0, 2048.0, 1024.0, 0.1, 650.0, (0.0, 0.0, 0.0), 0.0, 0.0)]
    for k in range(60):
        fixed.append((float(rng.uniform(50, 90)), float(rng.integers(256, 4096)), float(rng.integers(256, 4096)), 0.1,
                      float(rng.uniform(100, 5000)), tuple(float(x) for x in rng.uniform(-1000, 1000, 3)),
                      float(np.float32(rng.uniform(-6.3, 6.3))), float(np.float32(rng.uniform(-1.5, 1.5)))))
    for fov_deg, w, h, zn, zf, pos, yaw, pitch in fixed:
        out = np.zeros(256, np.uint8)
        p = np.array(pos, np.float32)
        fov = ref.ref_radians(fov_deg)
        ref.ref_setup_camera(fov, w, h, zn, zf, p.ctypes.data, yaw, pitch, out.ctypes.data)
        f = np.frombuffer(out.tobytes(), np.float32)[:46].copy()
        f[43] = f[44] = 0.0  # pyramidWidth/Height: not written by SetupCamera
        cases.append({"fov_deg": fov_deg, "fov_rad_bits": bits([fov])[0], "w": w, "h": h, "znear": zn, "zfar": zf,
                      "pos": list(pos), "yaw": yaw, "pitch": pitch, "view_data_bits": bits(f)})
    ops = {"mat4_inverse": [], "quat_to_mat4": [], "quat_from_angle_axis": [], "mat4_mul_vec4": []}
    for _ in range(40):
        m = rng.uniform(-3, 3, 16).astype(np.float32)
        m[0] += 4
        m[5] += 4
        m[10] += 4
        m[15] = 1.0
        r = np.zeros(16, np.float32)
        ref.ref_mat4_inverse(m.ctypes.data, r.ctypes.data)
        ops["mat4_inverse"].append({"in": bits(m), "out": bits(r)})
        q = rng.normal(size=4).astype(np.float32)
        ref.ref_quat_to_mat4(q.ctypes.data, r.ctypes.data)
        ops["quat_to_mat4"].append({"in": bits(q), "out": bits(r)})
        ax = rng.normal(size=3).astype(np.float32)
        ang = np.float32(rng.uniform(-7, 7))
        qo = np.zeros(4, np.float32)
        ref.ref_quat_from_angle_axis(ax.ctypes.data, float(ang), qo.ctypes.data)
        ops["quat_from_angle_axis"].append({"axis": bits(ax), "angle": bits([ang])[0], "out": bits(qo)})
        v = rng.uniform(-100, 100, 4).astype(np.float32)
        vo = np.zeros(4, np.float32)
        ref.ref_mat4_mul_vec4(m.ctypes.data, v.ctypes.data, vo.ctypes.data)
        ops["mat4_mul_vec4"].append({"m": bits(m), "v": bits(v), "out": bits(vo)})
    pyr = [{"w": int(w), "h": int(h), "prev_pow2": [int(ref.ref_previous_pow2(int(w))), int(ref.ref_previous_pow2(int(h)))],
            "mips": int(ref.ref_mip_levels(int(ref.ref_previous_pow2(int(w))), int(ref.ref_previous_pow2(int(h)))))}
           for w, h in [(1920, 1080), (1280, 720), (512, 512), (1024, 1024), (2048, 1024), (3840, 2160), (1, 1), (3, 7)]]
    groups = [{"n": n, "groups": int(ref.ref_group_size(n, 64))} for n in (0, 1, 63, 64, 65, 4194304)]
    OUT.write_text(json.dumps({"source": "reference camera + BlitML compiled from its sources (oracle/_ref)",
                               "cameras": cases, "ops": ops, "pyramid_extent": pyr, "group_size": groups}, indent=0))
    print(f"wrote {OUT} ({len(cases)} cameras)")


if __name__ == "__main__":
    main()
