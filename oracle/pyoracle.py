"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of oracle/_ref (the reference's compiled camera code).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_camera.so"

_lib = None
_ref = None


class OrcPyramid(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("base_w", C.c_uint32), ("base_h", C.c_uint32), ("mips", C.c_uint32),
                ("level_offset", C.c_uint32 * 32)]


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        lib = C.CDLL(str(LIB))
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        lib.orc_cull.restype = C.c_int64
        lib.orc_cull.argtypes = [vp, u32, vp, u32, vp, u32, vp, u32, vp, vp, u32, u32, u32, u32, vp, vp, vp, u32, vp,
                                 i32]
        lib.orc_pyramid_build.restype = None
        lib.orc_pyramid_build.argtypes = [vp, u32, u32, u32, u32, u32, vp, i32]
        lib.orc_pyramid_texels.restype = C.c_uint64
        lib.orc_pyramid_texels.argtypes = [u32, u32, u32, vp]
        lib.orc_pyramid_init.restype = None
        lib.orc_pyramid_init.argtypes = [C.POINTER(OrcPyramid), vp, u32, u32, u32]
        lib.orc_rotate_quat.argtypes = [vp, vp, vp]
        lib.orc_sphere_to_view.argtypes = [vp, f32, vp, f32, vp, vp, vp, vp]
        lib.orc_frustum.restype = i32
        lib.orc_frustum.argtypes = [vp, f32, f32, f32, f32, f32, f32, f32]
        lib.orc_lod_select.restype = u32
        lib.orc_lod_select.argtypes = [vp, f32, f32, f32, u32, u32, vp]
        lib.orc_project_sphere.restype = i32
        lib.orc_project_sphere.argtypes = [vp, f32, f32, f32, f32, vp]
        lib.orc_hiz_level.restype = u32
        lib.orc_hiz_level.argtypes = [f32, u32]
        lib.orc_footprint_min.restype = f32
        lib.orc_footprint_min.argtypes = [vp, u32, u32, f32, f32]
        lib.orc_occlusion.restype = i32
        lib.orc_occlusion.argtypes = [vp, C.POINTER(OrcPyramid), f32, f32, vp, f32, f32]
        lib.orc_camera_setup.restype = None
        lib.orc_camera_setup.argtypes = [f32, f32, f32, f32, f32, vp, f32, f32, vp]
        lib.orc_radians.restype = f32
        lib.orc_radians.argtypes = [f32]
        lib.orc_mat4_inverse.argtypes = [vp, vp]
        lib.orc_quat_to_mat4.argtypes = [vp, vp]
        lib.orc_quat_from_angle_axis.argtypes = [vp, f32, vp]
        lib.orc_previous_pow2.restype = u32
        lib.orc_previous_pow2.argtypes = [u32]
        lib.orc_mip_levels.restype = u32
        lib.orc_mip_levels.argtypes = [u32, u32]
        lib.orc_cluster_expand.restype = C.c_int64
        lib.orc_cluster_expand.argtypes = [vp, u32, vp, u32, vp, u32, vp, u32, vp, vp, vp, C.c_uint64]
        lib.orc_cluster_cull.restype = C.c_int64
        lib.orc_cluster_cull.argtypes = [vp, C.c_uint64, vp, vp, vp, vp, vp, u32, vp, C.c_uint64]
        lib.orc_cluster_visible.restype = i32
        lib.orc_cluster_visible.argtypes = [vp, vp, vp, vp, u32]
        lib.orc_threads_available.restype = i32
        lib.orc_cull_instanced.restype = C.c_int64
        lib.orc_cull_instanced.argtypes = [vp, u32, vp, u32, vp, u32, vp, u32, vp, u32, vp, vp, vp, u32]
        _lib = lib
    return _lib


def load_ref():
    """The reference's own camera/BlitML code compiled into oracle/_ref (None when not built)."""
    global _ref
    if _ref is None and REF_LIB.exists():
        lib = C.CDLL(str(REF_LIB))
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        lib.ref_setup_camera.argtypes = [f32, f32, f32, f32, f32, vp, f32, f32, vp]
        lib.ref_radians.restype = f32
        lib.ref_radians.argtypes = [f32]
        for n in ("ref_mat4_inverse", "ref_quat_to_mat4", "ref_normalize_plane"):
            getattr(lib, n).argtypes = [vp, vp]
        lib.ref_mat4_mul.argtypes = [vp, vp, vp]
        lib.ref_mat4_mul_vec4.argtypes = [vp, vp, vp]
        lib.ref_cross.argtypes = [vp, vp, vp]
        lib.ref_quat_from_angle_axis.argtypes = [vp, f32, vp]
        lib.ref_length3.restype = f32
        lib.ref_length3.argtypes = [vp]
        for n in ("ref_previous_pow2", "ref_group_size"):
            getattr(lib, n).restype = u32
        lib.ref_previous_pow2.argtypes = [u32]
        lib.ref_group_size.argtypes = [u32, u32]
        lib.ref_mip_levels.restype = u32
        lib.ref_mip_levels.argtypes = [u32, u32]
        _ref = lib
    return _ref


def p(a):
    return None if a is None else a.ctypes.data


def pyramid(texels: np.ndarray, base_w: int, base_h: int, mips: int) -> OrcPyramid:
    pr = OrcPyramid()
    load().orc_pyramid_init(C.byref(pr), p(texels), base_w, base_h, mips)
    pr._keep = texels
    return pr


def cull(scene, view_bytes: bytes, pass_: int, pyr=None, flags: int = 0, begin: int = 0, count: int = 0,
         vis_in=None, vis_out=None, capacity: int | None = None, threads: int = 0, want_lod: bool = False):
    """Runs one pass. scene = (objects, transforms, surfaces, lods) numpy structured arrays.
    Returns (cmds (n, 6) uint32 view as structured INDIRECT_DRAW bytes, count[, lods])."""
    objects, transforms, surfaces, lods = scene
    n_total = count if (begin or count) else len(objects)
    cap = n_total if capacity is None else capacity
    cmds = np.zeros(max(cap, 1) * 6, np.uint32)
    lod_out = np.zeros(max(cap, 1), np.uint32) if want_lod else None
    vb = np.frombuffer(view_bytes, np.uint8).copy()
    n = load().orc_cull(p(objects), len(objects), p(transforms), len(transforms), p(surfaces), len(surfaces), p(lods),
                        len(lods), p(vb), C.byref(pyr) if pyr is not None else None, pass_, flags, begin, count,
                        p(vis_in), p(vis_out), p(cmds), cap, p(lod_out), threads)
    if n < 0:
        raise ValueError("orc_cull rejected its arguments")
    k = min(n, cap)
    out = cmds[:k * 6].reshape(k, 6)
    if want_lod:
        return out, int(n), lod_out[:k]
    return out, int(n)


def pyramid_build(depth: np.ndarray, base_w: int, base_h: int, mips: int, threads: int = 0) -> np.ndarray:
    h, w = depth.shape
    n = load().orc_pyramid_texels(base_w, base_h, mips, None)
    out = np.zeros(n, np.float32)
    d = np.ascontiguousarray(depth, np.float32)
    load().orc_pyramid_build(p(d), w, h, base_w, base_h, mips, p(out), threads)
    return out


def camera(fov_deg: float, w: float, h: float, znear: float, zfar: float, pos=(0.0, 0.0, 0.0), yaw: float = 0.0,
           pitch: float = 0.0) -> bytes:
    lib = load()
    out = np.zeros(256, np.uint8)
    pp = np.array(pos, np.float32)
    lib.orc_camera_setup(lib.orc_radians(fov_deg), w, h, znear, zfar, p(pp), yaw, pitch, p(out))
    return out.tobytes()


def camera_rad(fov_rad: float, w: float, h: float, znear: float, zfar: float, pos=(0.0, 0.0, 0.0), yaw: float = 0.0,
               pitch: float = 0.0) -> bytes:
    out = np.zeros(256, np.uint8)
    pp = np.array(pos, np.float32)
    load().orc_camera_setup(fov_rad, w, h, znear, zfar, p(pp), yaw, pitch, p(out))
    return out.tobytes()


def view_floats(view_bytes: bytes) -> np.ndarray:
    return np.frombuffer(view_bytes, np.float32)


def cluster_expand(scene, view_bytes: bytes, pyr=None, capacity: int = 1 << 24):
    objects, transforms, surfaces, lods = scene
    out = np.zeros(capacity * 3, np.uint32)
    vb = np.frombuffer(view_bytes, np.uint8).copy()
    n = load().orc_cluster_expand(p(objects), len(objects), p(transforms), len(transforms), p(surfaces), len(surfaces),
                                  p(lods), len(lods), p(vb), C.byref(pyr) if pyr is not None else None, p(out), capacity)
    k = min(n, capacity)
    return out[:k * 3].reshape(k, 3), int(n)


def cluster_cull(recs: np.ndarray, objects, transforms, clusters, view_bytes: bytes, pyr, tests: int,
                 capacity: int | None = None):
    recs = np.ascontiguousarray(recs, np.uint32)
    n_recs = recs.shape[0]
    cap = n_recs if capacity is None else capacity
    cmds = np.zeros(max(cap, 1) * 6, np.uint32)
    vb = np.frombuffer(view_bytes, np.uint8).copy()
    n = load().orc_cluster_cull(p(recs), n_recs, p(objects), p(transforms), p(clusters), p(vb),
                                C.byref(pyr) if pyr is not None else None, tests, p(cmds), cap)
    k = min(n, cap)
    return cmds[:k * 6].reshape(k, 6), int(n)


def cull_instanced(scene, view_bytes: bytes, per_lod: int, capacity: int | None = None):
    """Returns (instance_ids (n_lods, per_lod) uint32 with unwritten slots = 0xFFFFFFFF, counters (n_lods, 2),
    commands (k, 8) uint32, number of commands)."""
    objects, transforms, surfaces, lods = scene
    nl = len(lods)
    ids = np.full(nl * per_lod, 0xFFFFFFFF, np.uint32)
    counters = np.zeros(nl * 2, np.uint32)
    cap = nl if capacity is None else capacity
    cmds = np.zeros(max(cap, 1) * 8, np.uint32)
    vb = np.frombuffer(view_bytes, np.uint8).copy()
    n = load().orc_cull_instanced(p(objects), len(objects), p(transforms), len(transforms), p(surfaces),
                                  len(surfaces), p(lods), nl, p(vb), per_lod, p(ids), p(counters), p(cmds), cap)
    k = min(n, cap)
    return ids.reshape(nl, per_lod), counters.reshape(nl, 2), cmds[:k * 8].reshape(k, 8), int(n)
