"""Jobs of more than 4096 tiles, whose emit takes its output offsets from a scan of the tile counts instead of adding
the earlier tiles' counts in every workgroup (linear in the object count), bit-exact against the oracle:
- one view of 17 000 000 objects: the PLAIN pass, and the two-pass frame with the frame cache (more than 4096 tiles
  in the EARLY pass and more than 4096 super-tiles in the cached LATE pass);
- a batch whose largest job has 5 000 000 objects next to small jobs (the batch's global prefix);
- the cluster cull of a 4 500 000-object scene (scanned record bases and survivor offsets).
Needs a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = bz.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def pyramid(ctx):
    depth = H.c1_depth()
    buf = ctx.alloc(depth.nbytes)
    buf.upload(depth)
    p = ctx.pyramid_create(1024, 512, 11)
    ctx.pyramid_build(p, buf.ptr, 1920, 1080)
    ctx.check_kernels()
    yield p, orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    ctx.pyramid_destroy(p)
    buf.free()


def read(ctx, cmds, cnt, cap):
    out, n = ctx.read_commands(cmds, cnt, capacity=cap)
    return H.cmds_u32(out), n


def test_17m_objects_one_view(ctx, pyramid):
    n = 17000000
    assert n // 1024 > 4096 and n // 4096 > 4096  # scanned: decide tiles and cached-LATE super-tiles
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    p, opyr = pyramid
    v = H.view(pyr=(1024, 512))
    vb = H.view_bytes(v)
    cap = n // 4
    cmds, cnt = ctx.alloc(cap * 24), ctx.alloc(4)
    ctx.cull(h, v, A.PASS_PLAIN, cmds, cnt, capacity=cap)
    ctx.check_kernels()
    got = read(ctx, cmds, cnt, cap)
    exp = orc.cull(scene, vb, A.PASS_PLAIN, capacity=cap, threads=16)
    assert got[1] == exp[1] and got[1] > 1000000
    np.testing.assert_array_equal(got[0], exp[0])
    vis0 = bz.synth_visibility(H.SEED_C1, n, 0.5)
    ctx.set_visibility(h, vis0)
    ctx.cull(h, v, A.PASS_EARLY, cmds, cnt, capacity=cap, flags=A.CULL_FRAME_CACHE)
    ctx.check_kernels()
    got = read(ctx, cmds, cnt, cap)
    exp = orc.cull(scene, vb, A.PASS_EARLY, vis_in=vis0, capacity=cap, threads=16)
    assert got[1] == exp[1]
    np.testing.assert_array_equal(got[0], exp[0])
    ctx.cull(h, v, A.PASS_LATE, cmds, cnt, capacity=cap, pyramid=p, flags=A.CULL_FRAME_CACHE)
    ctx.check_kernels()
    got = read(ctx, cmds, cnt, cap)
    vexp = vis0.copy()
    exp = orc.cull(scene, vb, A.PASS_LATE, opyr, vis_in=vexp, vis_out=vexp, capacity=cap, threads=16)
    assert got[1] == exp[1] and got[1] > 0
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    assert ctx.frame_cache_stats(h) == (1, 1)
    cmds.free()
    cnt.free()
    ctx.scene_destroy(h)


def test_batch_with_a_5m_object_job(ctx, pyramid):
    n = 5000000
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    p, opyr = pyramid
    ranges = [(0, 3000), (3000, 5000000 - 3000 - 70000), (5000000 - 70000, 70000)]  # small, > 4096 tiles, small
    views = [H.view(pyr=(1024, 512), yaw=0.2 * k) for k in range(len(ranges))]
    bufs, jobs = [], []
    for (b, c), v in zip(ranges, views):
        cmds, cnt = ctx.alloc(c * 24), ctx.alloc(4)
        bufs.append((cmds, cnt, c))
        jobs.append(ctx.view_job(h, v, A.PASS_TEMPORAL, cmds, cnt, capacity=c, pyramid=p, begin=b, n=c))
    ctx.cull_batch(jobs)
    ctx.check_kernels()
    for (b, c), v, (cmds, cnt, cap) in zip(ranges, views, bufs):
        got = read(ctx, cmds, cnt, cap)
        exp = orc.cull(scene, H.view_bytes(v), A.PASS_TEMPORAL, opyr, begin=b, count=c, threads=16)
        assert got[1] == exp[1]
        np.testing.assert_array_equal(got[0], exp[0])
        cmds.free()
        cnt.free()
    ctx.scene_destroy(h)


def test_cluster_cull_4_5m_objects(ctx, pyramid):
    n = 4500000
    sc = bz.synth_scene(H.SEED_C4 + 7, n, 128, 8, clusters=(16, 128))
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    ctx.scene_set_clusters(h, sc.clusters)
    p, opyr = pyramid
    v = H.view(pyr=(1024, 512))
    tests = A.CLUSTER_FRUSTUM | A.CLUSTER_HIZ | A.CLUSTER_CONE
    dcap = cap = 1 << 21
    disp, dcnt, cmds, cnt = ctx.alloc(dcap * 12), ctx.alloc(4), ctx.alloc(cap * 24), ctx.alloc(4)
    ctx.cull_clusters(h, v, tests, disp, dcnt, cmds, cnt, pyramid=p, dispatch_capacity=dcap, capacity=cap)
    ctx.check_kernels()
    nd = int(dcnt.download(np.uint32, 1)[0])
    scene = (sc.objects, sc.transforms, sc.surfaces, sc.lods)
    vb = H.view_bytes(v)
    recs, e_nd = orc.cluster_expand(scene, vb, None, capacity=dcap)
    e_cmds, e_nc = orc.cluster_cull(recs, sc.objects, sc.transforms, sc.clusters, vb, opyr, tests, capacity=cap)
    assert nd == e_nd and nd > 500000
    got_r = disp.download(np.uint32, nd * 3).reshape(-1, 3)
    np.testing.assert_array_equal(got_r, np.ascontiguousarray(recs).view(np.uint32).reshape(-1, 3))
    got_c, nc = read(ctx, cmds, cnt, cap)
    assert nc == e_nc
    np.testing.assert_array_equal(got_c, e_cmds)
    for b in (disp, dcnt, cmds, cnt):
        b.free()
    ctx.scene_destroy(h)
