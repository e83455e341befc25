"""Python host API over libbzcull (thin: numpy arrays in, device buffers and handles out).

Everything here forwards to the C ABI in include/bzcull.h; no cull arithmetic is done in Python.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _capi as A
from ._capi import (CLUSTER, CLUSTER_DISPATCH, INDIRECT_DRAW, LOD_DATA, MESH_TRANSFORM, PRIMITIVE_SURFACE,
                    RENDER_OBJECT, BzError, CameraViewData)


def lib():
    return A.load()


def status_str(s: int) -> str:
    return lib().bz_status_str(s).decode()


class DeviceBuffer:
    """Device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.check(lib().bz_dev_alloc(ctx.handle, self.nbytes, C.byref(p)), "bz_dev_alloc")
        self.ptr = p.value

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes, (a.nbytes, self.nbytes)
        self.ctx.check(lib().bz_memcpy_h2d(self.ctx.handle, self.ptr, A.ptr(a), a.nbytes, stream), "bz_memcpy_h2d")

    def download(self, dtype, count: int, stream=None, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            assert offset_bytes + out.nbytes <= self.nbytes
            self.ctx.check(lib().bz_memcpy_d2h(self.ctx.handle, A.ptr(out), self.ptr + offset_bytes, out.nbytes,
                                               stream), "bz_memcpy_d2h")
        return out

    def fill_u32(self, value: int, count: int | None = None, stream=None):
        n = self.nbytes // 4 if count is None else count
        self.ctx.check(lib().bz_memset_d32(self.ctx.handle, self.ptr, value, n, stream), "bz_memset_d32")

    def free(self):
        if self.ptr:
            lib().bz_dev_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.free()
        except Exception:
            pass


class MappedBuffer(DeviceBuffer):
    """Device pointer onto imported external memory; unmapped by Context.extmem_release, not by free()."""

    def __init__(self, ctx: "Context", ptr: int, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = ptr

    def free(self):
        self.ptr = None


class Context:
    def __init__(self, device: int = 0):
        cfg = A.Config(device=device)
        h = C.c_void_p()
        s = lib().bz_create(C.byref(cfg), C.byref(h))
        if s != A.OK:
            raise BzError(s, f"bz_create: {status_str(s)}")
        self.handle = h.value
        self._buffers: list[DeviceBuffer] = []

    # ------------------------------------------------------------------ basics
    def check(self, s: int, what: str):
        if s != A.OK:
            msg = lib().bz_last_error(self.handle).decode() if self.handle else ""
            raise BzError(s, f"{what}: {status_str(s)}: {msg}")

    def close(self):
        if self.handle:
            for b in self._buffers:
                b.free()
            self._buffers.clear()
            lib().bz_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def alloc(self, nbytes: int) -> DeviceBuffer:
        b = DeviceBuffer(self, nbytes)
        self._buffers.append(b)
        return b

    def sync(self, stream=None):
        self.check(lib().bz_stream_sync(self.handle, stream), "bz_stream_sync")

    def check_kernels(self, stream=None):
        self.check(lib().bz_check(self.handle, stream), "bz_check")

    def set_wait_limit(self, sleeps: int):
        """Sleeps a kernel's bounded wait may take before giving up (pyramid builds launched afterwards)."""
        self.check(lib().bz_set_wait_limit(self.handle, sleeps), "bz_set_wait_limit")

    def stream(self):
        s = C.c_void_p()
        self.check(lib().bz_stream_create(self.handle, C.byref(s)), "bz_stream_create")
        return s.value

    def device_info(self):
        name = C.create_string_buffer(64)
        ncu = C.c_uint32()
        hbm = C.c_uint64()
        self.check(lib().bz_device_info(self.handle, name, 64, C.byref(ncu), C.byref(hbm)), "bz_device_info")
        return name.value.decode(), ncu.value, hbm.value

    # ------------------------------------------------------------------ scenes
    def scene_create(self, objects, transforms, surfaces, lods) -> int:
        objects = np.ascontiguousarray(objects, dtype=RENDER_OBJECT)
        transforms = np.ascontiguousarray(transforms, dtype=MESH_TRANSFORM)
        surfaces = np.ascontiguousarray(surfaces, dtype=PRIMITIVE_SURFACE)
        lods = np.ascontiguousarray(lods, dtype=LOD_DATA)
        h = C.c_uint64()
        self.check(lib().bz_scene_create(self.handle, A.ptr(objects), len(objects), A.ptr(transforms), len(transforms),
                                         A.ptr(surfaces), len(surfaces), A.ptr(lods), len(lods), C.byref(h)),
                   "bz_scene_create")
        return h.value

    def scene_set_clusters(self, scene: int, clusters):
        clusters = np.ascontiguousarray(clusters, dtype=CLUSTER)
        self.check(lib().bz_scene_set_clusters(self.handle, scene, A.ptr(clusters), len(clusters)),
                   "bz_scene_set_clusters")

    def scene_destroy(self, scene: int):
        self.check(lib().bz_scene_destroy(self.handle, scene), "bz_scene_destroy")

    def set_visibility(self, scene: int, vis: np.ndarray, stream=None):
        vis = np.ascontiguousarray(vis, dtype=np.uint32)
        self.check(lib().bz_scene_set_visibility(self.handle, scene, A.ptr(vis), len(vis), stream),
                   "bz_scene_set_visibility")

    def get_visibility(self, scene: int, n: int, stream=None) -> np.ndarray:
        out = np.empty(n, dtype=np.uint32)
        self.check(lib().bz_scene_get_visibility(self.handle, scene, A.ptr(out), n, stream), "bz_scene_get_visibility")
        return out

    def visibility_ptr(self, scene: int) -> int:
        p = C.c_void_p()
        self.check(lib().bz_scene_visibility_ptr(self.handle, scene, C.byref(p)), "bz_scene_visibility_ptr")
        return p.value

    def get_transforms(self, scene: int, first: int, n: int, stream=None) -> np.ndarray:
        """The resident transforms [first, first + n) in the reference's MeshTransform layout."""
        out = np.empty(n, dtype=MESH_TRANSFORM)
        self.check(lib().bz_scene_get_transforms(self.handle, scene, first, n, A.ptr(out), stream),
                   "bz_scene_get_transforms")
        return out

    def frame_cache_stats(self, scene: int) -> tuple[int, int]:
        """(EARLY passes that recorded survivors, LATE passes that read them) for the scene."""
        rec, reu = C.c_uint64(), C.c_uint64()
        self.check(lib().bz_scene_frame_cache_stats(self.handle, scene, C.byref(rec), C.byref(reu)),
                   "bz_scene_frame_cache_stats")
        return rec.value, reu.value

    def update_transforms(self, scene: int, ids: np.ndarray, transforms: np.ndarray, stream=None):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        transforms = np.ascontiguousarray(transforms, dtype=MESH_TRANSFORM)
        self.check(lib().bz_scene_update_transforms(self.handle, scene, A.ptr(ids), A.ptr(transforms), len(ids),
                                                    stream), "bz_scene_update_transforms")

    def transform_staging(self, scene: int, capacity: int):
        """Pinned staging slot as numpy views (ids uint32[capacity], transforms MESH_TRANSFORM[capacity]); fill the
        first n entries, then commit_transforms(scene, n)."""
        ids_p, xf_p = C.c_void_p(), C.c_void_p()
        self.check(lib().bz_scene_transform_staging(self.handle, scene, capacity, C.byref(ids_p), C.byref(xf_p)),
                   "bz_scene_transform_staging")
        ids = np.ctypeslib.as_array((C.c_uint32 * capacity).from_address(ids_p.value))
        xf = np.frombuffer((C.c_uint8 * (capacity * MESH_TRANSFORM.itemsize)).from_address(xf_p.value),
                           dtype=MESH_TRANSFORM)
        return ids, xf

    def commit_transforms(self, scene: int, n: int, stream=None):
        self.check(lib().bz_scene_commit_transforms(self.handle, scene, n, stream), "bz_scene_commit_transforms")

    # ------------------------------------------------------------------ pyramids
    def pyramid_create(self, base_w: int, base_h: int, mips: int) -> int:
        h = C.c_uint64()
        self.check(lib().bz_pyramid_create(self.handle, base_w, base_h, mips, C.byref(h)), "bz_pyramid_create")
        return h.value

    def pyramid_build(self, pyramid: int, depth_dev: int, w: int, h: int, stream=None):
        self.check(lib().bz_pyramid_build(self.handle, pyramid, depth_dev, w, h, stream), "bz_pyramid_build")

    def pyramid_read(self, pyramid: int, mips: int) -> tuple[np.ndarray, np.ndarray]:
        tex = C.c_void_p()
        n = C.c_uint64()
        offs = (C.c_uint32 * mips)()
        self.check(lib().bz_pyramid_storage(self.handle, pyramid, C.byref(tex), C.byref(n), offs), "bz_pyramid_storage")
        out = np.empty(n.value, dtype=np.float32)
        self.check(lib().bz_memcpy_d2h(self.handle, A.ptr(out), tex.value, out.nbytes, None), "bz_memcpy_d2h")
        return out, np.array(list(offs), dtype=np.uint32)

    def pyramid_destroy(self, pyramid: int):
        self.check(lib().bz_pyramid_destroy(self.handle, pyramid), "bz_pyramid_destroy")

    # ------------------------------------------------------------------ cull
    def cull(self, scene: int, view: CameraViewData, pass_: int, cmds: DeviceBuffer, count: DeviceBuffer,
             capacity: int | None = None, pyramid: int = 0, flags: int = 0, begin: int = 0, n: int = 0,
             vis_in: int | None = None, vis_out: int | None = None, stream=None):
        if capacity is None:
            capacity = cmds.nbytes // 24
        a = A.CullArgs(scene=scene, pyramid=pyramid, pass_=pass_, flags=flags, view=C.pointer(view),
                       cmds_dev=cmds.ptr, count_dev=count.ptr, capacity=capacity, object_begin=begin, object_count=n,
                       visibility_in_dev=vis_in, visibility_out_dev=vis_out)
        self.check(lib().bz_cull(self.handle, C.byref(a), stream), "bz_cull")

    def cull_clusters(self, scene: int, view: CameraViewData, tests: int, dispatch: DeviceBuffer,
                      dispatch_count: DeviceBuffer, cmds: DeviceBuffer, count: DeviceBuffer, pyramid: int = 0,
                      dispatch_capacity: int | None = None, capacity: int | None = None, flags: int = 0,
                      stream=None):
        if dispatch_capacity is None:
            dispatch_capacity = dispatch.nbytes // 12
        if capacity is None:
            capacity = cmds.nbytes // 24
        a = A.ClusterCullArgs(scene=scene, pyramid=pyramid, cluster_tests=tests, flags=flags, view=C.pointer(view),
                              dispatch_dev=dispatch.ptr, dispatch_count_dev=dispatch_count.ptr,
                              dispatch_capacity=dispatch_capacity, cmds_dev=cmds.ptr, count_dev=count.ptr,
                              capacity=capacity)
        self.check(lib().bz_cull_clusters(self.handle, C.byref(a), stream), "bz_cull_clusters")

    def cull_instanced(self, scene: int, view: CameraViewData, per_lod: int, ids: DeviceBuffer, counters: DeviceBuffer,
                       cmds: DeviceBuffer, count: DeviceBuffer, capacity: int | None = None, stream=None):
        if capacity is None:
            capacity = cmds.nbytes // 32
        a = A.InstancedCullArgs(scene=scene, view=C.pointer(view), instances_per_lod=per_lod, capacity=capacity,
                                instance_ids_dev=ids.ptr, counters_dev=counters.ptr, cmds_dev=cmds.ptr,
                                count_dev=count.ptr)
        self.check(lib().bz_cull_instanced(self.handle, C.byref(a), stream), "bz_cull_instanced")

    # ------------------------------------------------------------------ interop (Vulkan-exported fds)
    def extmem_import(self, fd: int, size: int) -> int:
        h = C.c_uint64()
        self.check(lib().bz_extmem_import(self.handle, fd, size, C.byref(h)), "bz_extmem_import")
        return h.value

    def extmem_map(self, mem: int, offset: int, size: int) -> MappedBuffer:
        p = C.c_void_p()
        self.check(lib().bz_extmem_map(self.handle, mem, offset, size, C.byref(p)), "bz_extmem_map")
        return MappedBuffer(self, p.value, size)

    def extmem_release(self, mem: int):
        self.check(lib().bz_extmem_release(self.handle, mem), "bz_extmem_release")

    def extsem_import(self, fd: int, kind: int = A.EXTSEM_BINARY) -> int:
        h = C.c_uint64()
        self.check(lib().bz_extsem_import(self.handle, fd, kind, C.byref(h)), "bz_extsem_import")
        return h.value

    def extsem_wait(self, sem: int, value: int = 0, stream=None):
        self.check(lib().bz_extsem_wait(self.handle, sem, value, stream), "bz_extsem_wait")

    def extsem_signal(self, sem: int, value: int = 0, stream=None):
        self.check(lib().bz_extsem_signal(self.handle, sem, value, stream), "bz_extsem_signal")

    def extsem_release(self, sem: int):
        self.check(lib().bz_extsem_release(self.handle, sem), "bz_extsem_release")

    def read_commands(self, cmds: DeviceBuffer, count: DeviceBuffer, capacity: int | None = None, stream=None):
        n = int(count.download(np.uint32, 1, stream)[0])
        cap = cmds.nbytes // 24 if capacity is None else capacity
        return cmds.download(INDIRECT_DRAW, min(n, cap), stream), n

    def set_tuning(self, kernel: int, block: int, items: int, grid_per_cu: int = 0):
        t = A.Tuning(block=block, items=items, grid_per_cu=grid_per_cu)
        self.check(lib().bz_set_tuning(self.handle, kernel, C.byref(t)), "bz_set_tuning")

    def autotune(self, kernel: int, jobs: list[A.ViewJob], cache_path: str | None = None, stream=None) -> A.Tuning:
        arr = (A.ViewJob * len(jobs))(*jobs)
        t = A.Tuning()
        self.check(lib().bz_autotune(self.handle, kernel, arr, len(jobs),
                                     cache_path.encode() if cache_path else None, C.byref(t), stream), "bz_autotune")
        return t

    def get_tuning(self, kernel: int) -> A.Tuning:
        t = A.Tuning()
        self.check(lib().bz_get_tuning(self.handle, kernel, C.byref(t)), "bz_get_tuning")
        return t

    def cull_batch(self, jobs: list[A.ViewJob], stream=None):
        arr = (A.ViewJob * len(jobs))(*jobs)
        self.check(lib().bz_cull_batch(self.handle, arr, len(jobs), stream), "bz_cull_batch")

    @staticmethod
    def view_job(scene: int, view: CameraViewData, pass_: int, cmds: DeviceBuffer, count: DeviceBuffer,
                 capacity: int | None = None, pyramid: int = 0, begin: int = 0, n: int = 0, vis_in: int | None = None,
                 vis_out: int | None = None) -> A.ViewJob:
        if capacity is None:
            capacity = cmds.nbytes // 24
        return A.ViewJob(scene=scene, pyramid=pyramid, pass_=pass_, object_begin=begin, object_count=n,
                         capacity=capacity, view=view, visibility_in_dev=vis_in, visibility_out_dev=vis_out,
                         cmds_dev=cmds.ptr, count_dev=count.ptr)

    def pyramid_build_batch(self, jobs: list[tuple[int, int, int, int]], stream=None):
        """jobs: (pyramid, depth_dev, w, h) each."""
        arr = (A.PyramidJob * len(jobs))(*[A.PyramidJob(pyramid=p, depth_dev=d, w=w, h=h) for p, d, w, h in jobs])
        self.check(lib().bz_pyramid_build_batch(self.handle, arr, len(jobs), stream), "bz_pyramid_build_batch")

    def batch_graph_create(self, pyr_jobs: list[tuple[int, int, int, int]], jobs: list[A.ViewJob]) -> int:
        parr = (A.PyramidJob * max(len(pyr_jobs), 1))(*[A.PyramidJob(pyramid=p, depth_dev=d, w=w, h=h)
                                                         for p, d, w, h in pyr_jobs])
        jarr = (A.ViewJob * max(len(jobs), 1))(*jobs)
        g = C.c_uint64()
        self.check(lib().bz_batch_graph_create(self.handle, parr, len(pyr_jobs), jarr, len(jobs), C.byref(g)),
                   "bz_batch_graph_create")
        return g.value

    def graph_launch(self, graph: int, stream=None):
        self.check(lib().bz_graph_launch(self.handle, graph, stream), "bz_graph_launch")

    def graph_destroy(self, graph: int):
        self.check(lib().bz_graph_destroy(self.handle, graph), "bz_graph_destroy")

    # ------------------------------------------------------------------ serving
    def server_start(self, budget_bytes: int, inflight: int = 4, max_batch: int = 32, flags: int = 0):
        cfg = A.ServerCfg(hbm_budget_bytes=budget_bytes, max_inflight_batches=inflight, max_batch_requests=max_batch,
                          flags=flags)
        self.check(lib().bz_server_start(self.handle, C.byref(cfg)), "bz_server_start")

    def server_stop(self):
        self.check(lib().bz_server_stop(self.handle), "bz_server_stop")

    @staticmethod
    def request(scene: int, view: CameraViewData, depth_seed: int, depth_w: int, depth_h: int, occluders: int,
                pass_: int = A.PASS_TEMPORAL, deadline_ns: int = 0, user_tag: int = 0) -> A.Request:
        return A.Request(scene=scene, view=view, depth_seed=depth_seed, depth_w=depth_w, depth_h=depth_h,
                         occluders=occluders, pass_=pass_, deadline_ns=deadline_ns, user_tag=user_tag)

    def submit(self, req: A.Request) -> int:
        t = C.c_uint64()
        self.check(lib().bz_server_submit(self.handle, C.byref(req), C.byref(t)), "bz_server_submit")
        return t.value

    def submit_schedule(self, reqs: list[A.Request], offsets_ns) -> tuple[np.ndarray, np.ndarray]:
        n = len(reqs)
        arr = (A.Request * n)(*reqs)
        offs = np.ascontiguousarray(offsets_ns, dtype=np.uint64)
        tickets = np.zeros(n, np.uint64)
        st = np.zeros(n, np.int32)
        self.check(lib().bz_server_submit_schedule(self.handle, arr, offs.ctypes.data_as(C.POINTER(C.c_uint64)), n,
                                                   tickets.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                   st.ctypes.data_as(C.POINTER(C.c_int32))), "bz_server_submit_schedule")
        return tickets, st

    def poll(self, ticket: int):
        r = A.Result()
        s = lib().bz_server_poll(self.handle, ticket, C.byref(r))
        if s == A.E_NOT_READY:
            return None
        self.check(s, "bz_server_poll")
        return r

    def wait(self, ticket: int, timeout_s: float = 0.0) -> A.Result:
        r = A.Result()
        self.check(lib().bz_server_wait(self.handle, ticket, int(timeout_s * 1e9), C.byref(r)), "bz_server_wait")
        return r

    def fetch(self, ticket: int, capacity: int):
        out = np.zeros(max(capacity, 1), INDIRECT_DRAW)
        n = C.c_uint32()
        self.check(lib().bz_server_fetch(self.handle, ticket, A.ptr(out), capacity, C.byref(n)), "bz_server_fetch")
        return out[:min(n.value, capacity)], n.value

    def release(self, ticket: int):
        self.check(lib().bz_server_release(self.handle, ticket), "bz_server_release")

    def cancel(self, ticket: int):
        self.check(lib().bz_server_cancel(self.handle, ticket), "bz_server_cancel")

    def server_stats(self) -> A.ServerStats:
        st = A.ServerStats()
        self.check(lib().bz_server_get_stats(self.handle, C.byref(st)), "bz_server_get_stats")
        return st

    def synth_depth_device(self, seed: int, w: int, h: int, occluders: int, p11: float, znear: float, depth_dev: int,
                           stream=None):
        self.check(lib().bz_synth_depth_device(self.handle, seed, w, h, occluders, p11, znear, depth_dev, stream),
                   "bz_synth_depth_device")


# ---------------------------------------------------------------------------------------------------- helpers
def camera_setup(fov_deg: float, width: float, height: float, znear: float = 0.1, draw_distance: float = 650.0,
                 position=(0.0, 0.0, 0.0), yaw: float = 0.0, pitch: float = 0.0) -> CameraViewData:
    """SetupCamera (blitzenCamera.cpp:14-41) for a field of view in degrees (BlitML::Radians conversion)."""
    v = CameraViewData()
    fov = np.float32(fov_deg) * (np.float32(math.pi) / np.float32(180.0))
    pos = (C.c_float * 3)(*position)
    s = lib().bz_camera_setup(float(np.float32(fov)), width, height, znear, draw_distance, pos, yaw, pitch, C.byref(v))
    if s != A.OK:
        raise BzError(s, "bz_camera_setup")
    return v


def view_bytes(v: CameraViewData) -> bytes:
    return bytes(memoryview(v))


@dataclass
class SynthScene:
    objects: np.ndarray
    transforms: np.ndarray
    surfaces: np.ndarray
    lods: np.ndarray
    clusters: np.ndarray | None = None


def synth_scene(seed: int, n_objects: int, n_surfaces: int, lods_per_surface: int, pos=(-500.0, 500.0),
                scale=(0.5, 2.0), clusters=(0, 0)) -> SynthScene:
    cfg = A.SynthSceneCfg(seed=seed, n_objects=n_objects, n_surfaces=n_surfaces, lods_per_surface=lods_per_surface,
                          clusters_min=clusters[0], clusters_max=clusters[1], pos_min=pos[0], pos_max=pos[1],
                          scale_min=scale[0], scale_max=scale[1])
    objects = np.zeros(n_objects, RENDER_OBJECT)
    transforms = np.zeros(n_objects, MESH_TRANSFORM)
    surfaces = np.zeros(n_surfaces, PRIMITIVE_SURFACE)
    lods = np.zeros(n_surfaces * lods_per_surface, LOD_DATA)
    s = lib().bz_synth_scene(C.byref(cfg), A.ptr(objects), A.ptr(transforms), A.ptr(surfaces), A.ptr(lods))
    if s != A.OK:
        raise BzError(s, "bz_synth_scene")
    cl = None
    if clusters[0]:
        n = C.c_uint32()
        s = lib().bz_synth_clusters(C.byref(cfg), A.ptr(lods), None, C.byref(n))
        if s != A.OK:
            raise BzError(s, "bz_synth_clusters(count)")
        cl = np.zeros(n.value, CLUSTER)
        s = lib().bz_synth_clusters(C.byref(cfg), A.ptr(lods), A.ptr(cl), C.byref(n))
        if s != A.OK:
            raise BzError(s, "bz_synth_clusters")
    return SynthScene(objects, transforms, surfaces, lods, cl)


def synth_depth(seed: int, w: int, h: int, occluders: int, p11: float, znear: float) -> np.ndarray:
    out = np.empty((h, w), np.float32)
    s = lib().bz_synth_depth(seed, w, h, occluders, p11, znear, A.ptr(out))
    if s != A.OK:
        raise BzError(s, "bz_synth_depth")
    return out


def synth_visibility(seed: int, n: int, p: float = 0.5) -> np.ndarray:
    out = np.empty(n, np.uint32)
    s = lib().bz_synth_visibility(seed, n, p, A.ptr(out))
    if s != A.OK:
        raise BzError(s, "bz_synth_visibility")
    return out


def synth_transform_updates(seed: int, frame: int, base: np.ndarray, count: int):
    ids = np.empty(count, np.uint32)
    out = np.empty(count, MESH_TRANSFORM)
    base = np.ascontiguousarray(base, dtype=MESH_TRANSFORM)
    s = lib().bz_synth_transform_updates(seed, frame, A.ptr(base), len(base), count, A.ptr(ids), A.ptr(out))
    if s != A.OK:
        raise BzError(s, "bz_synth_transform_updates")
    return ids, out


def reference_pyramid_extent(w: int, h: int):
    bw, bh, m = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib().bz_pyramid_reference_extent(w, h, C.byref(bw), C.byref(bh), C.byref(m))
    return bw.value, bh.value, m.value


def pyramid_level_offsets(base_w: int, base_h: int, mips: int):
    offs, dims, off = [], [], 0
    for L in range(mips):
        w, h = max(1, base_w >> L), max(1, base_h >> L)
        offs.append(off)
        dims.append((w, h))
        off += w * h
    return offs, dims, off
