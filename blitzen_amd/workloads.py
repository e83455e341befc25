"""Seeded synthetic workloads of the benchmark harness and tests.

batched views: 256 views of uneven object counts (log-uniform in [1 024, 1 048 576], about 32 M object-view pairs)
    over a shared 256-surface x 8-LOD table; each view has its own seeded camera (fov U[50, 90] deg) and its own
    depth buffer (512x512, 1024x1024 or 2048x1024, 64-1024 occluder rectangles) reduced to a min-depth pyramid;
    late-pass semantics per view (frustum + Hi-Z, newly visible objects emitted, visibility written).
served requests: 8 resident scenes (262 144 / 524 288 / 1 M / 1 M / 2 M / 2 M / 4 M / 4 M objects) and Poisson
    arrivals of cull requests, each with a seeded camera and a depth buffer of 512x512 to 1920x1080.

Every view's objects are its own range of one large scene (no view reads another's objects), so a batch streams
all of its object-view pairs from HBM like separate frames would.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _capi as A
from . import cull as bz

BATCH_SEED = 0xB1172E02
SERVE_SEED = 0xB1172E03
DEPTH_SIZES = [(512, 512), (1024, 1024), (2048, 1024)]
SERVE_SCENES = [262144, 524288, 1048576, 1048576, 2097152, 2097152, 4194304, 4194304]


@dataclass
class ViewPlan:
    begin: int
    count: int
    fov: float
    w: int
    h: int
    occluders: int
    pos: tuple
    yaw: float
    pitch: float
    depth_seed: int


def batched_plan(n_views: int = 256, seed: int = BATCH_SEED, min_objects: int = 1024,
                 max_objects: int = 1048576) -> list[ViewPlan]:
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(min_objects), math.log(max_objects), n_views)).astype(np.int64)
    sizes = rng.integers(0, len(DEPTH_SIZES), n_views)
    out, off = [], 0
    for v in range(n_views):
        w, h = DEPTH_SIZES[int(sizes[v])]
        out.append(ViewPlan(begin=off, count=int(counts[v]), fov=float(rng.uniform(50, 90)), w=w, h=h,
                            occluders=int(rng.integers(64, 1025)), pos=tuple(float(x) for x in rng.uniform(-100, 100, 3)),
                            yaw=float(rng.uniform(-math.pi, math.pi)), pitch=float(rng.uniform(-0.5, 0.5)),
                            depth_seed=seed * 1000 + v))
        off += int(counts[v])
    return out


class _Ptr:
    """A device address with the attributes the Context helpers read (ptr, nbytes)."""

    def __init__(self, ptr: int, nbytes: int = 0):
        self.ptr = ptr
        self.nbytes = nbytes


@dataclass
class BatchedViews:
    """Device-resident batched-views workload: one scene holding every view's objects, per-view depth buffers,
    pyramids, commands and counts, and the job tables for the batched and the per-view forms."""
    ctx: bz.Context
    plan: list
    scene: int
    n_objects: int
    views: list = field(default_factory=list)
    pyramids: list = field(default_factory=list)
    depth: list = field(default_factory=list)
    buffers: list = field(default_factory=list)
    jobs: list = field(default_factory=list)
    pyr_jobs: list = field(default_factory=list)
    cull_args: list = field(default_factory=list)

    @classmethod
    def build(cls, ctx: bz.Context, plan: list[ViewPlan], seed: int = BATCH_SEED, stream=None) -> "BatchedViews":
        total = sum(p.count for p in plan)
        sc = bz.synth_scene(seed, total, 256, 8)
        scene = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
        del sc
        self = cls(ctx=ctx, plan=plan, scene=scene, n_objects=total)
        vis0 = bz.synth_visibility(seed, total, 0.5)
        vin = ctx.alloc(total * 4)
        vin.upload(vis0)
        vout = ctx.alloc(total * 4)
        cmds = ctx.alloc(total * 24)
        counts = ctx.alloc(len(plan) * 4)
        self.buffers += [vin, vout, cmds, counts]
        for k, p in enumerate(plan):
            v = bz.camera_setup(p.fov, p.w, p.h, 0.1, 650.0, p.pos, p.yaw, p.pitch)
            bw, bh, mips = bz.reference_pyramid_extent(p.w, p.h)
            v.pyramidWidth, v.pyramidHeight = float(bw), float(bh)
            d = ctx.alloc(p.w * p.h * 4)
            ctx.synth_depth_device(p.depth_seed, p.w, p.h, p.occluders, v.proj5, 0.1, d.ptr, stream)
            pyr = ctx.pyramid_create(bw, bh, mips)
            self.views.append(v)
            self.depth.append(d)
            self.pyramids.append(pyr)
            cmd_ptr = cmds.ptr + p.begin * 24
            cnt_ptr = counts.ptr + 4 * k
            self.jobs.append(A.ViewJob(scene=scene, pyramid=pyr, pass_=A.PASS_LATE, object_begin=p.begin,
                                       object_count=p.count, capacity=p.count, view=v, visibility_in_dev=vin.ptr,
                                       visibility_out_dev=vout.ptr, cmds_dev=cmd_ptr, count_dev=cnt_ptr))
            self.pyr_jobs.append((pyr, d.ptr, p.w, p.h))
            self.cull_args.append(A.CullArgs(scene=scene, pyramid=pyr, pass_=A.PASS_LATE, flags=0, view=C.pointer(v),
                                             cmds_dev=cmd_ptr, count_dev=cnt_ptr, capacity=p.count,
                                             object_begin=p.begin, object_count=p.count,
                                             visibility_in_dev=vin.ptr, visibility_out_dev=vout.ptr))
        self.vis0 = vis0
        self.counts_buf = counts
        ctx.sync(stream)
        return self

    def run_per_view(self, stream=None):
        """The plain form: per view, one pyramid build and one cull call (three kernel launches)."""
        lib, h = bz.lib(), self.ctx.handle
        for k in range(len(self.plan)):
            p = self.plan[k]
            s = lib.bz_pyramid_build(h, self.pyramids[k], self.depth[k].ptr, p.w, p.h, stream)
            if s == A.OK:
                s = lib.bz_cull(h, C.byref(self.cull_args[k]), stream)
            if s != A.OK:
                self.ctx.check(s, "per-view cull")

    def counts(self, stream=None) -> np.ndarray:
        return self.counts_buf.download(np.uint32, len(self.plan), stream)

    def free(self):
        for g in self.pyramids:
            self.ctx.pyramid_destroy(g)
        for b in self.depth + self.buffers:
            b.free()
        self.ctx.scene_destroy(self.scene)


def algorithmic_bytes_batched(plan: list[ViewPlan], commands: int) -> int:
    """Per view: N*48 (objects, transforms, visibility read + write) + the depth buffer read + the pyramid written;
    plus 24 B per command and 4 B per count."""
    total = 0
    for p in plan:
        bw, bh, mips = bz.reference_pyramid_extent(p.w, p.h)
        texels = sum(max(1, bw >> L) * max(1, bh >> L) for L in range(mips))
        total += p.count * 48 + p.w * p.h * 4 + texels * 4 + 4
    return total + commands * 24


SERVE_DEPTH_SIZES = [(512, 512), (800, 600), (1024, 768), (1280, 720), (1600, 900), (1920, 1080)]
SERVE_OCCLUDERS = 128


def served_scenes(ctx: bz.Context, seed: int = SERVE_SEED, sizes=SERVE_SCENES):
    """The resident scenes of the serving workload (positions U[-500,500]^3, scales U[0.5,2], 256 surfaces x 8 LODs).
    Returns [(handle, n_objects)]."""
    out = []
    for k, n in enumerate(sizes):
        sc = bz.synth_scene(seed + k, n, 256, 8)
        out.append((ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods), n))
        del sc
    return out


def served_requests(scenes, rate: float, duration_s: float, seed: int = SERVE_SEED):
    """Poisson arrivals at `rate` req/s for `duration_s`: each request picks a resident scene, a seeded camera
    (fov U[50, 90] deg) and a depth buffer of 512x512 .. 1920x1080 with 128 occluders (late-pass Hi-Z against its
    own pyramid). Returns (requests, arrival offsets in ns, objects per request)."""
    rng = np.random.default_rng([seed, int(rate)])
    n = max(1, int(rate * duration_s))
    offsets = np.cumsum(rng.exponential(1e9 / rate, n)).astype(np.uint64)
    reqs, objs = [], []
    for i in range(n):
        h, n_obj = scenes[int(rng.integers(len(scenes)))]
        w, hh = SERVE_DEPTH_SIZES[int(rng.integers(len(SERVE_DEPTH_SIZES)))]
        v = bz.camera_setup(float(rng.uniform(50, 90)), w, hh, 0.1, 650.0, tuple(rng.uniform(-50, 50, 3)),
                            float(rng.uniform(-math.pi, math.pi)), float(rng.uniform(-0.5, 0.5)))
        reqs.append(bz.Context.request(h, v, depth_seed=seed * 1000003 + int(rate) * 7 + i, depth_w=w, depth_h=hh,
                                       occluders=SERVE_OCCLUDERS, pass_=A.PASS_TEMPORAL, user_tag=i))
        objs.append(n_obj)
    return reqs, offsets, np.array(objs, np.int64)


CLUSTER_SEED = 0xB1172E04


@dataclass
class ClusterFrames:
    """Cluster-mode workload: 1 048 576 objects over 128 surfaces with 16-128 clusters of 64 triangles each (cone
    axis uniform on the sphere, cutoff U[-0.5, 0.9]); per frame 1% of the transforms get seeded rotation updates,
    then the two-level cull runs: the object pass of PreClusterDrawCull.comp.glsl:13-47 (frustum + LOD, no Hi-Z),
    the expansion of the chosen LODs' clusters, and the cluster frustum + Hi-Z + cone tests against a 1920x1080
    depth buffer's 1024x512 11-mip pyramid."""
    ctx: bz.Context
    scene: int
    n_objects: int
    transforms: np.ndarray
    view: object
    pyramid: int
    buffers: list

    @classmethod
    def build(cls, ctx: bz.Context, n_objects: int = 1048576, seed: int = CLUSTER_SEED, stream=None):
        sc = bz.synth_scene(seed, n_objects, 128, 8, clusters=(16, 128))
        h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
        ctx.scene_set_clusters(h, sc.clusters)
        v = bz.camera_setup(70.0, 1920, 1080, 0.1, 650.0)
        v.pyramidWidth, v.pyramidHeight = 1024.0, 512.0
        depth = ctx.alloc(1920 * 1080 * 4)
        ctx.synth_depth_device(seed, 1920, 1080, 512, v.proj5, 0.1, depth.ptr, stream)
        pyr = ctx.pyramid_create(1024, 512, 11)
        ctx.pyramid_build(pyr, depth.ptr, 1920, 1080, stream)
        max_records = n_objects * int(sc.lods["clusterCount"].max())
        disp = ctx.alloc(min(max_records, 1 << 27) * 12)
        cmds = ctx.alloc(min(max_records, 1 << 27) * 24)
        counts = ctx.alloc(16)
        self = cls(ctx=ctx, scene=h, n_objects=n_objects, transforms=sc.transforms.copy(), view=v, pyramid=pyr,
                   buffers=[depth, disp, cmds, counts])
        self.disp, self.cmds, self.counts = disp, cmds, counts
        return self

    def updates(self, frame: int):
        """The frame's seeded rotation updates (1% of the transforms): (ids, new transforms)."""
        return bz.synth_transform_updates(CLUSTER_SEED, frame, self.transforms, self.n_objects // 100)

    def frame(self, ids, new, stream=None):
        self.ctx.update_transforms(self.scene, ids, new, stream)
        self.cull(stream)

    def cull(self, stream=None):
        """The cluster cull alone (object pass, expansion, cluster tests, emit)."""
        cnt = self.counts
        self.ctx.cull_clusters(self.scene, self.view, A.CLUSTER_FRUSTUM | A.CLUSTER_HIZ | A.CLUSTER_CONE, self.disp,
                               _Ptr(cnt.ptr), self.cmds, _Ptr(cnt.ptr + 4), pyramid=self.pyramid, stream=stream)

    def result_counts(self, stream=None):
        c = self.counts.download(np.uint32, 2, stream)
        return int(c[0]), int(c[1])

    def free(self):
        self.ctx.pyramid_destroy(self.pyramid)
        for b in self.buffers:
            b.free()
        self.ctx.scene_destroy(self.scene)
