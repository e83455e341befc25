"""Parity at the benchmarked sizes, bit-exact against the oracle on the same seeded inputs:
- the frame of 4 194 304 objects (1920x1080 depth, 1024x512 11-mip pyramid, last-frame visibility Bernoulli(0.5)):
  EARLY + LATE with and without the frame cache, the PLAIN pass and the LOD-instanced output;
- the 256 batched views (38 250 177 object-view pairs, per-view 512x512 / 1024x1024 / 2048x1024 depth and
  reference-extent pyramids) replayed from one graph;
- the cluster mode scene of 1 048 576 objects (128 surfaces, 16-128 clusters each) after 1% transform updates;
- the serving loop over its 8 resident scenes (15 466 496 objects) under 160 Poisson arrivals at 32 000 req/s.
Needs a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from blitzen_amd import workloads as W
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu
N_FRAME = 4194304


@pytest.fixture(scope="module")
def ctx():
    c = bz.Context(0)
    yield c
    c.close()


def read(ctx, cmds, cnt, cap):
    out, n = ctx.read_commands(cmds, cnt, capacity=cap)
    return H.cmds_u32(out), n


@pytest.fixture(scope="module")
def frame_scene(ctx):
    sc = H.c1_scene(N_FRAME)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    depth = H.c1_depth()
    dbuf = ctx.alloc(depth.nbytes)
    dbuf.upload(depth)
    p = ctx.pyramid_create(1024, 512, 11)
    ctx.pyramid_build(p, dbuf.ptr, 1920, 1080)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    yield h, scene, p, opyr
    ctx.pyramid_destroy(p)
    dbuf.free()
    ctx.scene_destroy(h)


@pytest.mark.parametrize("flags", [A.CULL_FRAME_CACHE, 0])
def test_frame_4m_objects(ctx, frame_scene, flags):
    h, scene, p, opyr = frame_scene
    n = N_FRAME
    v = H.view(pyr=(1024, 512))
    vb = H.view_bytes(v)
    vis0 = bz.synth_visibility(H.SEED_C1, n, 0.5)
    ctx.set_visibility(h, vis0)
    cmds, cnt = ctx.alloc(n * 24), ctx.alloc(4)
    ctx.cull(h, v, A.PASS_EARLY, cmds, cnt, capacity=n, flags=flags)
    got_e = read(ctx, cmds, cnt, n)
    ctx.cull(h, v, A.PASS_LATE, cmds, cnt, capacity=n, pyramid=p, flags=flags)
    ctx.check_kernels()
    got_l = read(ctx, cmds, cnt, n)
    exp_e, ne = orc.cull(scene, vb, A.PASS_EARLY, vis_in=vis0)
    vexp = vis0.copy()
    exp_l, nl = orc.cull(scene, vb, A.PASS_LATE, opyr, vis_in=vis0, vis_out=vexp)
    assert (got_e[1], got_l[1]) == (ne, nl) and ne > 100000 and nl > 1000
    np.testing.assert_array_equal(got_e[0], exp_e)
    np.testing.assert_array_equal(got_l[0], exp_l)
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    cmds.free()
    cnt.free()


def test_plain_and_instanced_4m_objects(ctx, frame_scene):
    h, scene, _, _ = frame_scene
    n = N_FRAME
    v = H.view()
    vb = H.view_bytes(v)
    cmds, cnt = ctx.alloc(n * 24), ctx.alloc(4)
    ctx.cull(h, v, A.PASS_PLAIN, cmds, cnt, capacity=n)
    ctx.check_kernels()
    got = read(ctx, cmds, cnt, n)
    exp, ne = orc.cull(scene, vb, A.PASS_PLAIN)
    assert got[1] == ne
    np.testing.assert_array_equal(got[0], exp)
    cmds.free()
    cnt.free()
    nl, per_lod = len(scene[3]), 4096
    ids, counters = ctx.alloc(nl * per_lod * 4), ctx.alloc(nl * 8)
    icmds, icnt = ctx.alloc(nl * 32), ctx.alloc(4)
    ids.fill_u32(0xFFFFFFFF)
    ctx.cull_instanced(h, v, per_lod, ids, counters, icmds, icnt)
    ctx.check_kernels()
    k = int(icnt.download(np.uint32, 1)[0])
    e_ids, e_ctr, e_cmds, ek = orc.cull_instanced(scene, vb, per_lod)
    assert k == ek
    np.testing.assert_array_equal(ids.download(np.uint32, nl * per_lod).reshape(nl, per_lod), e_ids)
    np.testing.assert_array_equal(counters.download(np.uint32, nl * 2).reshape(nl, 2), e_ctr)
    np.testing.assert_array_equal(icmds.download(np.uint32, k * 8).reshape(k, 8), e_cmds)
    for b in (ids, counters, icmds, icnt):
        b.free()


def test_batched_256_views_graph(ctx):
    plan = W.batched_plan()
    bv = W.BatchedViews.build(ctx, plan)
    try:
        g = ctx.batch_graph_create(bv.pyr_jobs, bv.jobs)
        ctx.graph_launch(g)
        ctx.graph_launch(g)  # a replay gives the same result (the inputs do not change)
        ctx.check_kernels()
        counts = bv.counts()
        cmds_all = bv.buffers[2]
        vout = bv.buffers[1].download(np.uint32, bv.n_objects)
        sc = bz.synth_scene(W.BATCH_SEED, bv.n_objects, 256, 8)
        scene = H.as_tuple(sc)
        vexp = np.zeros(bv.n_objects, np.uint32)
        for k, p in enumerate(plan):
            v = bv.views[k]
            bw, bh, mips = bz.reference_pyramid_extent(p.w, p.h)
            depth = bz.synth_depth(p.depth_seed, p.w, p.h, p.occluders, v.proj5, 0.1)
            opyr = orc.pyramid(orc.pyramid_build(depth, bw, bh, mips), bw, bh, mips)
            exp, ne = orc.cull(scene, H.view_bytes(v), A.PASS_LATE, opyr, begin=p.begin, count=p.count,
                               vis_in=bv.vis0, vis_out=vexp)
            assert int(counts[k]) == ne, k
            got = cmds_all.download(np.uint32, ne * 6, offset_bytes=p.begin * 24).reshape(-1, 6)
            np.testing.assert_array_equal(got, exp)
        np.testing.assert_array_equal(vout, vexp)
        ctx.graph_destroy(g)
    finally:
        bv.free()


def test_cluster_mode_1m_objects(ctx):
    cf = W.ClusterFrames.build(ctx)
    try:
        sc = bz.synth_scene(W.CLUSTER_SEED, cf.n_objects, 128, 8, clusters=(16, 128))
        xf = sc.transforms.copy()
        for f in range(2):
            ids, new = cf.updates(f)
            cf.transforms[ids] = new  # the workload's next frame rotates the updated transforms further
            xf[ids] = new
            cf.frame(ids, new)
        ctx.check_kernels()
        nd, nc = cf.result_counts()
        depth = bz.synth_depth(W.CLUSTER_SEED, 1920, 1080, 512, cf.view.proj5, 0.1)
        opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
        scene = (sc.objects, xf, sc.surfaces, sc.lods)
        vb = H.view_bytes(cf.view)
        # the object pass is PreClusterDrawCull's frustum + LOD (no object Hi-Z); the pyramid serves the cluster tests
        recs, e_nd = orc.cluster_expand(scene, vb, None, capacity=1 << 22)
        e_cmds, e_nc = orc.cluster_cull(recs, sc.objects, xf, sc.clusters, vb, opyr,
                                        A.CLUSTER_FRUSTUM | A.CLUSTER_HIZ | A.CLUSTER_CONE, capacity=1 << 22)
        assert (nd, nc) == (e_nd, e_nc) and nd > 100000
        got_r = cf.disp.download(np.uint32, nd * 3).reshape(-1, 3)
        np.testing.assert_array_equal(got_r, np.ascontiguousarray(recs).view(np.uint32).reshape(-1, 3))
        got_c = cf.cmds.download(np.uint32, nc * 6).reshape(-1, 6)
        np.testing.assert_array_equal(got_c, e_cmds)
    finally:
        cf.free()


def test_served_8_resident_scenes(ctx):
    """The serving workload's 8 resident scenes (15 466 496 objects) under a burst of Poisson arrivals at 32 000
    req/s: every request's commands equal the oracle's cull of its scene with its own device-rasterised depth."""
    import ctypes as C

    scenes = W.served_scenes(ctx)
    try:
        ctx.server_start(64 << 30, inflight=4, max_batch=64)
        try:
            reqs, offs, _ = W.served_requests(scenes, 32000, 0.005, seed=5)
            tickets, statuses = ctx.submit_schedule(reqs, offs)
            assert (statuses == A.OK).all()
            results = []
            for t, req in zip(tickets, reqs):
                r = ctx.wait(int(t), timeout_s=60)
                assert r.status == A.OK
                n_obj = dict((h, n) for h, n in scenes)[req.scene]
                cmds, n = ctx.fetch(int(t), n_obj)
                assert n == r.count
                results.append((req, H.cmds_u32(cmds)[:n], r.batch_size))
        finally:
            ctx.server_stop()
        assert max(b for _, _, b in results) > 1  # some requests were batched together
        host = {}
        for k, (h, n) in enumerate(scenes):
            host[h] = H.as_tuple(bz.synth_scene(W.SERVE_SEED + k, n, 256, 8))
        for req, got, _ in results:
            v = bz.CameraViewData()
            C.memmove(C.addressof(v), C.addressof(req.view), C.sizeof(v))
            bw, bh, mips = bz.reference_pyramid_extent(req.depth_w, req.depth_h)
            v.pyramidWidth, v.pyramidHeight = float(bw), float(bh)
            depth = bz.synth_depth(req.depth_seed, req.depth_w, req.depth_h, req.occluders, v.proj5, v.zNear)
            opyr = orc.pyramid(orc.pyramid_build(depth, bw, bh, mips), bw, bh, mips)
            exp, ne = orc.cull(host[req.scene], H.view_bytes(v), A.PASS_TEMPORAL, opyr)
            assert got.shape[0] == ne, req.user_tag
            np.testing.assert_array_equal(got, exp)
    finally:
        for h, _ in scenes:
            ctx.scene_destroy(h)
