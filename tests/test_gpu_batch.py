"""Batched multi-view culling and hipGraph replay vs the CPU oracle, bit-exact, through the C ABI. Needs a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

PYR_SHAPES = [(512, 512, 256, 256, 8), (1920, 1080, 1024, 512, 11), (2048, 1024, 1024, 512, 10),
              (1000, 700, 500, 350, 9), (8192, 4096, 4096, 2048, 12)]  # the last two take the level-by-level path


@pytest.fixture(scope="module")
def ctx():
    c = bz.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scenes(ctx):
    """Four scenes covering the kernel variants: LDS tables + identity transforms, small tables, tables too large for
    LDS, and an indirect (shared) transform stream."""
    out = []
    a = H.c1_scene(200000, seed=21)
    out.append(H.as_tuple(a))
    b = H.c0_scene()
    out.append(H.as_tuple(b))
    c = bz.synth_scene(23, 50000, 600, 4)  # 600 surfaces, 2400 LODs
    out.append(H.as_tuple(c))
    d = H.c1_scene(40000, seed=24)
    rng = np.random.default_rng(24)
    objs = d.objects.copy()
    objs["transformId"] = rng.integers(0, 20000, len(objs), dtype=np.uint32)
    out.append((objs, d.transforms[:20000].copy(), d.surfaces, d.lods))
    handles = [ctx.scene_create(*s) for s in out]
    yield list(zip(handles, out))
    for h in handles:
        ctx.scene_destroy(h)


def make_pyramids(ctx, seeds):
    """One pyramid per entry of PYR_SHAPES (cycled), depth synthesized on the device. Returns (jobs, info) where
    info = (handle, depth buffer, base_w, base_h, mips, w, h)."""
    jobs, info = [], []
    for k, seed in enumerate(seeds):
        w, h, bw, bh, mips = PYR_SHAPES[k % len(PYR_SHAPES)]
        buf = ctx.alloc(w * h * 4)
        ctx.synth_depth_device(seed, w, h, 64 + 61 * k, H.view(w=w, h=h).proj5, 0.1, buf.ptr)
        p = ctx.pyramid_create(bw, bh, mips)
        jobs.append((p, buf.ptr, w, h))
        info.append((p, buf, bw, bh, mips, w, h))
    return jobs, info


def oracle_pyramid(buf, bw, bh, mips, w, h):
    depth = buf.download(np.float32, w * h).reshape(h, w)
    return orc.pyramid(orc.pyramid_build(depth, bw, bh, mips), bw, bh, mips)


class Job:
    """One view's cull as both a bz_view_job and the oracle call that must give the same bytes."""

    def __init__(self, ctx, scene_h, scene, view, pass_, pyr=None, begin=0, count=0, capacity=None, seed=0):
        self.scene, self.view, self.pass_, self.pyr = scene, view, pass_, pyr
        self.begin, self.count = begin, count
        n = len(scene[0])
        sel = (n - begin) if count == 0 and begin == 0 else count
        self.cap = sel if capacity is None else capacity
        self.cmds = ctx.alloc(max(self.cap, 1) * 24)
        self.cnt = ctx.alloc(4)
        self.cnt.fill_u32(0xDEADBEEF, 1)
        self.vis_in = self.vis_out = None
        if pass_ == A.PASS_LATE:
            self.vis0 = bz.synth_visibility(seed, n, 0.5)
            self.vis_in = ctx.alloc(n * 4)
            self.vis_in.upload(self.vis0)
            self.vis_out = ctx.alloc(n * 4)
            self.vis_out.fill_u32(7)
        self.vj = ctx.view_job(scene_h, view, pass_, self.cmds, self.cnt, capacity=self.cap,
                               pyramid=pyr[0] if pyr else 0, begin=begin, n=count,
                               vis_in=self.vis_in.ptr if self.vis_in else None,
                               vis_out=self.vis_out.ptr if self.vis_out else None)

    def check(self, ctx, opyr):
        got, total = ctx.read_commands(self.cmds, self.cnt, capacity=self.cap)
        n = len(self.scene[0])
        vexp = None
        if self.pass_ == A.PASS_LATE:
            vexp = np.full(n, 7, np.uint32)
        exp, ne = orc.cull(self.scene, H.view_bytes(self.view), self.pass_, opyr, begin=self.begin, count=self.count,
                           vis_in=self.vis0 if self.pass_ == A.PASS_LATE else None, vis_out=vexp,
                           capacity=self.cap, threads=8)
        assert total == ne, (self.pass_, self.begin, self.count, total, ne)
        np.testing.assert_array_equal(H.cmds_u32(got), exp)
        if vexp is not None:
            np.testing.assert_array_equal(self.vis_out.download(np.uint32, n), vexp)

    def free(self):
        for b in (self.cmds, self.cnt, self.vis_in, self.vis_out):
            if b is not None:
                b.free()


def build_jobs(ctx, scenes, pinfo):
    rng = np.random.default_rng(5)
    jobs = []
    passes = [A.PASS_LATE, A.PASS_PLAIN, A.PASS_TEMPORAL]
    for k in range(14):
        sh, sc = scenes[k % len(scenes)]
        n = len(sc[0])
        p = pinfo[k % len(pinfo)]
        v = H.view(fov=float(rng.uniform(50, 90)), w=p[5], h=p[6], pos=tuple(rng.uniform(-50, 50, 3)),
                   yaw=float(rng.uniform(-3, 3)), pitch=float(rng.uniform(-0.5, 0.5)), pyr=(p[2], p[3]))
        pass_ = passes[k % 3]
        begin, count = (0, 0) if k % 4 == 0 else (int(rng.integers(0, n // 2)), int(rng.integers(1, n // 2)))
        cap = None if k % 5 else int(rng.integers(0, 50))
        jobs.append((Job(ctx, sh, sc, v, pass_, pyr=p if pass_ != A.PASS_PLAIN else None, begin=begin, count=count,
                         capacity=cap, seed=100 + k), k % len(pinfo)))
    sh, sc = scenes[1]
    jobs.append((Job(ctx, sh, sc, H.view(), A.PASS_PLAIN, begin=5, count=0), 0))            # empty job
    jobs.append((Job(ctx, sh, sc, H.view(), A.PASS_PLAIN, begin=len(sc[0]) - 1, count=1), 0))  # one object
    return jobs


def test_batch_matches_oracle(ctx, scenes):
    pjobs, pinfo = make_pyramids(ctx, [31, 32, 33, 34, 35])
    ctx.pyramid_build_batch(pjobs)
    ctx.check_kernels()
    opyrs = [oracle_pyramid(b, bw, bh, m, w, h) for (_, b, bw, bh, m, w, h) in pinfo]
    for k, (p, _, bw, bh, m, _, _) in enumerate(pinfo):
        got, _ = ctx.pyramid_read(p, m)
        np.testing.assert_array_equal(got.view(np.uint32), np.frombuffer(opyrs[k]._keep, np.uint32))
    jobs = build_jobs(ctx, scenes, pinfo)
    ctx.cull_batch([j.vj for j, _ in jobs])
    ctx.check_kernels()
    for j, pk in jobs:
        j.check(ctx, opyrs[pk] if j.pyr else None)
        j.free()
    for (p, b, *_) in pinfo:
        ctx.pyramid_destroy(p)
        b.free()


@pytest.mark.parametrize("mode", ["eager", "graph"])
@pytest.mark.parametrize("n_views", [600, 1100])
def test_batch_many_views(ctx, scenes, n_views, mode):
    """Batches past the job lookups' kernel-argument table (512 views): the first tiles staged in LDS (up to 1024
    views) and searched in global memory (beyond), each job a short range of its own; run eagerly, or captured in a
    graph and replayed twice (the counts poisoned before the second replay)."""
    sh, sc = scenes[0]
    n = len(sc[0])
    rng = np.random.default_rng(n_views)
    jobs = []
    for k in range(n_views):
        count = int(rng.integers(1, 2600))
        begin = int(rng.integers(0, n - count))
        jobs.append(Job(ctx, sh, sc, H.view(yaw=float(rng.uniform(-3, 3)), pitch=float(rng.uniform(-0.5, 0.5))),
                        A.PASS_PLAIN, begin=begin, count=count))
    try:
        if mode == "eager":
            ctx.cull_batch([j.vj for j in jobs])
            ctx.check_kernels()
            for j in jobs:
                j.check(ctx, None)
        else:
            g = ctx.batch_graph_create([], [j.vj for j in jobs])
            try:
                for rep in range(2):
                    if rep:
                        for j in jobs:
                            j.cnt.fill_u32(0xDEADBEEF, 1)
                        ctx.sync()
                    ctx.graph_launch(g)
                    ctx.check_kernels()
                    for j in jobs:
                        j.check(ctx, None)
            finally:
                ctx.graph_destroy(g)
    finally:
        for j in jobs:
            j.free()


@pytest.mark.parametrize("tuning", [(256, 2), (256, 4), (512, 2), (512, 4)])
def test_batch_tunings(ctx, scenes, tuning):
    old = ctx.get_tuning(1)
    ctx.set_tuning(1, *tuning)
    try:
        sh, sc = scenes[0]
        jobs = [Job(ctx, sh, sc, H.view(yaw=0.3 * k), A.PASS_PLAIN, begin=1000 * k, count=50000 + 977 * k)
                for k in range(6)]
        ctx.cull_batch([j.vj for j in jobs])
        ctx.check_kernels()
        for j in jobs:
            j.check(ctx, None)
            j.free()
    finally:
        ctx.set_tuning(1, old.block, old.items, old.grid_per_cu)


def test_all_empty_batch(ctx, scenes):
    sh, sc = scenes[1]
    jobs = [Job(ctx, sh, sc, H.view(), A.PASS_PLAIN, begin=7, count=0) for _ in range(3)]
    ctx.cull_batch([j.vj for j in jobs])
    ctx.check_kernels()
    for j in jobs:
        assert int(j.cnt.download(np.uint32, 1)[0]) == 0
        j.free()


def test_graph_replays(ctx, scenes):
    """Capture pyramid builds + batched cull once; replay with new depth and visibility contents written by eager
    work between replays; every replay must match the oracle on the inputs it saw."""
    pjobs, pinfo = make_pyramids(ctx, [41, 42, 43, 44, 45])
    jobs = build_jobs(ctx, scenes, pinfo)
    g = ctx.batch_graph_create(pjobs, [j.vj for j, _ in jobs])
    sh0, sc0 = scenes[1]
    probe = Job(ctx, sh0, sc0, H.view(yaw=1.0), A.PASS_PLAIN)
    try:
        for rep in range(4):
            if rep:
                for k, (p, buf, bw, bh, m, w, h) in enumerate(pinfo):
                    ctx.synth_depth_device(1000 * rep + k, w, h, 64 + 97 * k, H.view(w=w, h=h).proj5, 0.1, buf.ptr)
                for j, _ in jobs:
                    if j.pass_ == A.PASS_LATE:
                        j.vis0 = bz.synth_visibility(77 * rep + j.begin, len(j.scene[0]), 0.3 + 0.1 * rep)
                        j.vis_in.upload(j.vis0)
                        j.vis_out.fill_u32(7)
                    j.cnt.fill_u32(0xDEADBEEF, 1)
                ctx.sync()
            ctx.graph_launch(g)
            ctx.check_kernels()
            # eager work between replays, on the same stream
            ctx.cull(sh0, probe.view, A.PASS_PLAIN, probe.cmds, probe.cnt)
            ctx.check_kernels()
            probe.check(ctx, None)
            opyrs = [oracle_pyramid(b, bw, bh, m, w, h) for (_, b, bw, bh, m, w, h) in pinfo]
            for j, pk in jobs:
                j.check(ctx, opyrs[pk] if j.pyr else None)
    finally:
        ctx.graph_destroy(g)
        probe.free()
        for j, _ in jobs:
            j.free()
        for (p, b, *_) in pinfo:
            ctx.pyramid_destroy(p)
            b.free()


def test_batch_equals_single_view_calls(ctx, scenes):
    sh, sc = scenes[0]
    depth = H.c1_depth()
    buf = ctx.alloc(depth.nbytes)
    buf.upload(depth)
    p = ctx.pyramid_create(1024, 512, 11)
    ctx.pyramid_build(p, buf.ptr, 1920, 1080)
    views = [H.view(yaw=0.7 * k, pyr=(1024, 512)) for k in range(5)]
    batch = [Job(ctx, sh, sc, v, A.PASS_TEMPORAL, pyr=(p,)) for v in views]
    single = [Job(ctx, sh, sc, v, A.PASS_TEMPORAL, pyr=(p,)) for v in views]
    ctx.cull_batch([j.vj for j in batch])
    for j in single:
        ctx.cull(sh, j.view, A.PASS_TEMPORAL, j.cmds, j.cnt, pyramid=p)
    ctx.check_kernels()
    for a, b in zip(batch, single):
        ga, na = ctx.read_commands(a.cmds, a.cnt)
        gb, nb = ctx.read_commands(b.cmds, b.cnt)
        assert na == nb
        np.testing.assert_array_equal(H.cmds_u32(ga), H.cmds_u32(gb))
        a.free()
        b.free()
    ctx.pyramid_destroy(p)
    buf.free()


def test_single_view_large_tables(ctx, scenes):
    """Surface/LOD tables too large for LDS (600 surfaces, 2400 LODs) on the single-view path, every pass."""
    sh, sc = scenes[2]
    depth = H.c1_depth()
    buf = ctx.alloc(depth.nbytes)
    buf.upload(depth)
    p = ctx.pyramid_create(1024, 512, 11)
    ctx.pyramid_build(p, buf.ptr, 1920, 1080)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    v = H.view(pyr=(1024, 512))
    for pass_ in (A.PASS_PLAIN, A.PASS_TEMPORAL, A.PASS_LATE, A.PASS_ONPC):
        j = Job(ctx, sh, sc, v, pass_, pyr=(p,) if pass_ in (A.PASS_TEMPORAL, A.PASS_LATE) else None, seed=3)
        ctx.cull(sh, v, pass_, j.cmds, j.cnt, pyramid=p if j.pyr else 0,
                 vis_in=j.vis_in.ptr if j.vis_in else None, vis_out=j.vis_out.ptr if j.vis_out else None)
        ctx.check_kernels()
        j.check(ctx, opyr if j.pyr else None)
        j.free()
    ctx.pyramid_destroy(p)
    buf.free()


def test_autotune_and_cache(ctx, scenes, tmp_path):
    sh, sc = scenes[0]
    jobs = [Job(ctx, sh, sc, H.view(yaw=0.5 * k), A.PASS_PLAIN, begin=2000 * k, count=30000) for k in range(4)]
    old0, old1 = ctx.get_tuning(0), ctx.get_tuning(1)
    cache = tmp_path / "tune.json"
    try:
        t1 = ctx.autotune(1, [j.vj for j in jobs], cache_path=str(cache))
        t0 = ctx.autotune(0, [j.vj for j in jobs], cache_path=str(cache))
        assert (t1.block, t1.items) in [(256, 2), (256, 4), (512, 2), (512, 4)]
        assert t0.grid_per_cu in (0, 1, 2, 4)
        text = cache.read_text()
        assert "|k0|b16" in text and "|k1|b16" in text  # 120000 objects -> bucket 16
        again = ctx.autotune(1, [j.vj for j in jobs], cache_path=str(cache))  # served from the cache
        assert (again.block, again.items, again.grid_per_cu) == (t1.block, t1.items, t1.grid_per_cu)
        g = ctx.get_tuning(1)
        assert (g.block, g.items) == (t1.block, t1.items)
        ctx.cull_batch([j.vj for j in jobs])
        ctx.check_kernels()
        for j in jobs:
            j.check(ctx, None)
    finally:
        ctx.set_tuning(0, old0.block, old0.items, old0.grid_per_cu)
        ctx.set_tuning(1, old1.block, old1.items, old1.grid_per_cu)
        for j in jobs:
            j.free()
