"""ctypes binding of libbzcull.so (include/bzcull.h).

The library is loaded from this package directory only. When it is missing the import fails with an error naming
the build step: there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libbzcull.so"

# ---------------------------------------------------------------------------------------------------- layouts
RENDER_OBJECT = np.dtype([("transformId", "<u4"), ("surfaceId", "<u4")])
MESH_TRANSFORM = np.dtype([("pos", "<f4", 3), ("scale", "<f4"), ("orientation", "<f4", 4)])
PRIMITIVE_SURFACE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("materialId", "<u4"), ("lodOffset", "<u4"),
                              ("lodCount", "<u4"), ("vertexOffset", "<u4")])
LOD_DATA = np.dtype([("indexCount", "<u4"), ("firstIndex", "<u4"), ("clusterOffset", "<u4"), ("clusterCount", "<u4"),
                     ("error", "<f4"), ("padding", "<u4", 3)])
CLUSTER = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("coneAxis", "i1", 3), ("coneCutoff", "i1"),
                    ("dataOffset", "<u4"), ("vertexCount", "u1"), ("triangleCount", "u1"), ("padding", "u1", 2),
                    ("padding2", "<u4")])
INDIRECT_DRAW = np.dtype([("objectId", "<u4"), ("indexCount", "<u4"), ("instanceCount", "<u4"), ("firstIndex", "<u4"),
                          ("vertexOffset", "<i4"), ("firstInstance", "<u4")])
CLUSTER_DISPATCH = np.dtype([("objectId", "<u4"), ("lodIndex", "<u4"), ("clusterId", "<u4")])
INDIRECT_DRAW32 = np.dtype([("objId", "<u4"), ("indexCount", "<u4"), ("instanceCount", "<u4"), ("firstIndex", "<u4"),
                            ("vertexOffset", "<i4"), ("firstInstance", "<u4"), ("padding", "<u4", 2)])
LOD_INSTANCE_COUNTER = np.dtype([("instanceOffset", "<u4"), ("instanceCount", "<u4")])
for _dt, _sz in ((RENDER_OBJECT, 8), (MESH_TRANSFORM, 32), (PRIMITIVE_SURFACE, 32), (LOD_DATA, 32), (CLUSTER, 32),
                 (INDIRECT_DRAW, 24), (CLUSTER_DISPATCH, 12), (INDIRECT_DRAW32, 32), (LOD_INSTANCE_COUNTER, 8)):
    assert _dt.itemsize == _sz, (_dt, _sz)


class CameraViewData(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("projectionView", C.c_float * 16), ("position", C.c_float * 3),
                ("frustumRight", C.c_float), ("frustumLeft", C.c_float), ("frustumTop", C.c_float),
                ("frustumBottom", C.c_float), ("proj0", C.c_float), ("proj5", C.c_float), ("zNear", C.c_float),
                ("zFar", C.c_float), ("pyramidWidth", C.c_float), ("pyramidHeight", C.c_float),
                ("lodTarget", C.c_float), ("padding", C.c_uint8 * 72)]


assert C.sizeof(CameraViewData) == 256 and CameraViewData.lodTarget.offset == 180


class Config(C.Structure):
    _fields_ = [("device", C.c_int32), ("flags", C.c_uint32), ("reserved", C.c_uint64 * 6)]


class CullArgs(C.Structure):
    _fields_ = [("scene", C.c_uint64), ("pyramid", C.c_uint64), ("pass_", C.c_uint32), ("flags", C.c_uint32),
                ("view", C.POINTER(CameraViewData)), ("cmds_dev", C.c_void_p), ("count_dev", C.c_void_p),
                ("capacity", C.c_uint32), ("object_begin", C.c_uint32), ("object_count", C.c_uint32),
                ("visibility_in_dev", C.c_void_p), ("visibility_out_dev", C.c_void_p)]


class ClusterCullArgs(C.Structure):
    _fields_ = [("scene", C.c_uint64), ("pyramid", C.c_uint64), ("cluster_tests", C.c_uint32), ("flags", C.c_uint32),
                ("view", C.POINTER(CameraViewData)), ("dispatch_dev", C.c_void_p), ("dispatch_count_dev", C.c_void_p),
                ("dispatch_capacity", C.c_uint32), ("cmds_dev", C.c_void_p), ("count_dev", C.c_void_p),
                ("capacity", C.c_uint32)]


class InstancedCullArgs(C.Structure):
    _fields_ = [("scene", C.c_uint64), ("view", C.POINTER(CameraViewData)), ("instances_per_lod", C.c_uint32),
                ("capacity", C.c_uint32), ("instance_ids_dev", C.c_void_p), ("counters_dev", C.c_void_p),
                ("cmds_dev", C.c_void_p), ("count_dev", C.c_void_p)]


class ViewJob(C.Structure):
    _fields_ = [("scene", C.c_uint64), ("pyramid", C.c_uint64), ("pass_", C.c_uint32), ("object_begin", C.c_uint32),
                ("object_count", C.c_uint32), ("capacity", C.c_uint32), ("view", CameraViewData),
                ("visibility_in_dev", C.c_void_p), ("visibility_out_dev", C.c_void_p), ("cmds_dev", C.c_void_p),
                ("count_dev", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("block", C.c_uint32), ("items", C.c_uint32), ("grid_per_cu", C.c_uint32), ("reserved", C.c_uint32)]


class PyramidJob(C.Structure):
    _fields_ = [("pyramid", C.c_uint64), ("depth_dev", C.c_void_p), ("w", C.c_uint32), ("h", C.c_uint32)]


class ServerCfg(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_inflight_batches", C.c_uint32),
                ("max_batch_requests", C.c_uint32), ("flags", C.c_uint32), ("reserved0", C.c_uint32),
                ("reserved", C.c_uint64 * 4)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_uint64), ("view", CameraViewData), ("depth_seed", C.c_uint64), ("depth_w", C.c_uint32),
                ("depth_h", C.c_uint32), ("occluders", C.c_uint32), ("pass_", C.c_uint32),
                ("deadline_ns", C.c_uint64), ("user_tag", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("count", C.c_uint32), ("submit_ns", C.c_uint64), ("admit_ns", C.c_uint64),
                ("complete_ns", C.c_uint64), ("batch_id", C.c_uint64), ("batch_size", C.c_uint32),
                ("reserved", C.c_uint32), ("user_tag", C.c_uint64), ("arrival_ns", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cancelled", C.c_uint64), ("batches", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("budget_in_use", C.c_uint64), ("budget_peak", C.c_uint64), ("queue_depth", C.c_uint32),
                ("inflight_batches", C.c_uint32)]


class SynthSceneCfg(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_objects", C.c_uint32), ("n_surfaces", C.c_uint32),
                ("lods_per_surface", C.c_uint32), ("clusters_min", C.c_uint32), ("clusters_max", C.c_uint32),
                ("pos_min", C.c_float), ("pos_max", C.c_float), ("scale_min", C.c_float), ("scale_max", C.c_float)]


# ---------------------------------------------------------------------------------------------------- statuses
OK, E_INVALID, E_OOM, E_HIP, E_CAPACITY, E_BUDGET, E_TIMEOUT, E_NOT_READY, E_NOT_FOUND, E_STATE, E_NODEVICE, \
    E_INTERNAL, E_CANCELLED = 0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10, -11, -12
PASS_EARLY, PASS_LATE, PASS_PLAIN, PASS_TEMPORAL, PASS_ONPC = 0, 1, 2, 3, 4
DEFAULT_WAIT_LIMIT = 1048576
CULL_EARLY_HIZ = 0x1
CULL_FRAME_CACHE = 0x2
CLUSTER_FRUSTUM, CLUSTER_HIZ, CLUSTER_CONE = 0x1, 0x2, 0x4
CLUSTER_OBJECT_HIZ = 0x1  # bz_cluster_cull_args.flags
SERVER_RELEASE_ON_COMPLETE = 0x1
EXTSEM_BINARY = 0
EXTSEM_TIMELINE = 1

_VP = C.c_void_p
_U32, _U64, _F32, _I32 = C.c_uint32, C.c_uint64, C.c_float, C.c_int32
_PV = C.POINTER(CameraViewData)

# name -> (restype, argtypes); every function declared in include/bzcull.h
SIGNATURES = {
    "bz_abi_version": (_U32, []),
    "bz_create": (_I32, [C.POINTER(Config), C.POINTER(_VP)]),
    "bz_destroy": (None, [_VP]),
    "bz_status_str": (C.c_char_p, [_I32]),
    "bz_last_error": (C.c_char_p, [_VP]),
    "bz_device_info": (_I32, [_VP, C.c_char_p, _U32, C.POINTER(_U32), C.POINTER(_U64)]),
    "bz_scene_create": (_I32, [_VP, _VP, _U32, _VP, _U32, _VP, _U32, _VP, _U32, C.POINTER(_U64)]),
    "bz_scene_set_clusters": (_I32, [_VP, _U64, _VP, _U32]),
    "bz_scene_destroy": (_I32, [_VP, _U64]),
    "bz_scene_update_transforms": (_I32, [_VP, _U64, _VP, _VP, _U32, _VP]),
    "bz_scene_set_visibility": (_I32, [_VP, _U64, _VP, _U32, _VP]),
    "bz_scene_get_visibility": (_I32, [_VP, _U64, _VP, _U32, _VP]),
    "bz_scene_visibility_ptr": (_I32, [_VP, _U64, C.POINTER(_VP)]),
    "bz_scene_frame_cache_stats": (_I32, [_VP, _U64, C.POINTER(_U64), C.POINTER(_U64)]),
    "bz_scene_get_transforms": (_I32, [_VP, _U64, _U32, _U32, _VP, _VP]),
    "bz_scene_info": (_I32, [_VP, _U64, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32),
                             C.POINTER(_U64)]),
    "bz_pyramid_create": (_I32, [_VP, _U32, _U32, _U32, C.POINTER(_U64)]),
    "bz_pyramid_destroy": (_I32, [_VP, _U64]),
    "bz_pyramid_build": (_I32, [_VP, _U64, _VP, _U32, _U32, _VP]),
    "bz_pyramid_storage": (_I32, [_VP, _U64, C.POINTER(_VP), C.POINTER(_U64), C.POINTER(_U32)]),
    "bz_pyramid_reference_extent": (None, [_U32, _U32, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32)]),
    "bz_pyramid_build_batch": (_I32, [_VP, C.POINTER(PyramidJob), _U32, _VP]),
    "bz_cull": (_I32, [_VP, C.POINTER(CullArgs), _VP]),
    "bz_cull_noalloc": (_I32, [_VP, C.POINTER(CullArgs), _VP]),
    "bz_cull_instanced": (_I32, [_VP, C.POINTER(InstancedCullArgs), _VP]),
    "bz_scene_transform_staging": (_I32, [_VP, _U64, _U32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "bz_scene_commit_transforms": (_I32, [_VP, _U64, _U32, _VP]),
    "bz_extmem_import": (_I32, [_VP, C.c_int, C.c_uint64, C.POINTER(C.c_uint64)]),
    "bz_extmem_map": (_I32, [_VP, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p)]),
    "bz_extmem_release": (_I32, [_VP, C.c_uint64]),
    "bz_extsem_import": (_I32, [_VP, C.c_int, C.c_uint32, C.POINTER(C.c_uint64)]),
    "bz_extsem_wait": (_I32, [_VP, C.c_uint64, C.c_uint64, _VP]),
    "bz_extsem_signal": (_I32, [_VP, C.c_uint64, C.c_uint64, _VP]),
    "bz_extsem_release": (_I32, [_VP, C.c_uint64]),
    "bz_cull_clusters": (_I32, [_VP, C.POINTER(ClusterCullArgs), _VP]),
    "bz_cull_batch": (_I32, [_VP, C.POINTER(ViewJob), _U32, _VP]),
    "bz_check": (_I32, [_VP, _VP]),
    "bz_set_wait_limit": (_I32, [_VP, _U32]),
    "bz_get_tuning": (_I32, [_VP, _U32, C.POINTER(Tuning)]),
    "bz_set_tuning": (_I32, [_VP, _U32, C.POINTER(Tuning)]),
    "bz_autotune": (_I32, [_VP, _U32, C.POINTER(ViewJob), _U32, C.c_char_p, C.POINTER(Tuning), _VP]),
    "bz_batch_graph_create": (_I32, [_VP, C.POINTER(PyramidJob), _U32, C.POINTER(ViewJob), _U32, C.POINTER(_U64)]),
    "bz_graph_launch": (_I32, [_VP, _U64, _VP]),
    "bz_graph_destroy": (_I32, [_VP, _U64]),
    "bz_camera_setup": (_I32, [_F32, _F32, _F32, _F32, _F32, C.POINTER(_F32), _F32, _F32, _PV]),
    "bz_server_start": (_I32, [_VP, C.POINTER(ServerCfg)]),
    "bz_server_stop": (_I32, [_VP]),
    "bz_server_submit": (_I32, [_VP, C.POINTER(Request), C.POINTER(_U64)]),
    "bz_server_submit_schedule": (_I32, [_VP, C.POINTER(Request), C.POINTER(_U64), _U32, C.POINTER(_U64),
                                         C.POINTER(C.c_int32)]),
    "bz_server_poll": (_I32, [_VP, _U64, C.POINTER(Result)]),
    "bz_server_wait": (_I32, [_VP, _U64, _U64, C.POINTER(Result)]),
    "bz_server_fetch": (_I32, [_VP, _U64, _VP, _U32, C.POINTER(_U32)]),
    "bz_server_cancel": (_I32, [_VP, _U64]),
    "bz_server_release": (_I32, [_VP, _U64]),
    "bz_server_get_stats": (_I32, [_VP, C.POINTER(ServerStats)]),
    "bz_synth_scene": (_I32, [C.POINTER(SynthSceneCfg), _VP, _VP, _VP, _VP]),
    "bz_synth_clusters": (_I32, [C.POINTER(SynthSceneCfg), _VP, _VP, C.POINTER(_U32)]),
    "bz_synth_depth": (_I32, [_U64, _U32, _U32, _U32, _F32, _F32, _VP]),
    "bz_synth_depth_device": (_I32, [_VP, _U64, _U32, _U32, _U32, _F32, _F32, _VP, _VP]),
    "bz_synth_visibility": (_I32, [_U64, _U32, _F32, _VP]),
    "bz_synth_transform_updates": (_I32, [_U64, _U64, _VP, _U32, _U32, _VP, _VP]),
    "bz_dev_alloc": (_I32, [_VP, _U64, C.POINTER(_VP)]),
    "bz_dev_free": (_I32, [_VP, _VP]),
    "bz_memcpy_h2d": (_I32, [_VP, _VP, _VP, _U64, _VP]),
    "bz_memcpy_d2h": (_I32, [_VP, _VP, _VP, _U64, _VP]),
    "bz_memcpy_d2d": (_I32, [_VP, _VP, _VP, _U64, _VP]),
    "bz_memset_d32": (_I32, [_VP, _VP, _U32, _U64, _VP]),
    "bz_stream_create": (_I32, [_VP, C.POINTER(_VP)]),
    "bz_stream_destroy": (_I32, [_VP, _VP]),
    "bz_stream_sync": (_I32, [_VP, _VP]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libbzcull.so from the package directory (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m blitzen_amd.build` "
                              "(or __graft_entry__.build()); the cull path has no CPU fallback")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class BzError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to libbzcull must be C-contiguous"
    return a.ctypes.data
