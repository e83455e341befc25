"""Two-level cluster cull on the GPU vs the CPU oracle, bit-exact, through the C ABI. Needs a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

ALL = A.CLUSTER_FRUSTUM | A.CLUSTER_HIZ | A.CLUSTER_CONE


@pytest.fixture(scope="module")
def ctx():
    c = bz.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def cl_scene(ctx):
    sc = bz.synth_scene(H.SEED_C4, 60000, 128, 8, clusters=(16, 128))
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    ctx.scene_set_clusters(h, sc.clusters)
    yield h, sc
    ctx.scene_destroy(h)


@pytest.fixture(scope="module")
def pyramid(ctx):
    depth = H.c1_depth(seed=H.SEED_C4, occluders=512)
    buf = ctx.alloc(depth.nbytes)
    buf.upload(depth)
    p = ctx.pyramid_create(1024, 512, 11)
    ctx.pyramid_build(p, buf.ptr, 1920, 1080)
    ctx.check_kernels()
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    yield p, opyr
    ctx.pyramid_destroy(p)
    buf.free()


def run_gpu(ctx, h, v, tests, pyr=0, dcap=1 << 21, cap=1 << 21, flags=0):
    disp = ctx.alloc(max(dcap, 1) * 12)
    dcnt = ctx.alloc(4)
    cmds = ctx.alloc(max(cap, 1) * 24)
    cnt = ctx.alloc(4)
    dcnt.fill_u32(0xDEADBEEF, 1)
    cnt.fill_u32(0xDEADBEEF, 1)
    ctx.cull_clusters(h, v, tests, disp, dcnt, cmds, cnt, pyramid=pyr, dispatch_capacity=dcap, capacity=cap,
                      flags=flags)
    ctx.check_kernels()
    nd = int(dcnt.download(np.uint32, 1)[0])
    recs = disp.download(A.CLUSTER_DISPATCH, min(nd, dcap)) if min(nd, dcap) else np.zeros(0, A.CLUSTER_DISPATCH)
    out, n = ctx.read_commands(cmds, cnt, capacity=cap)
    for b in (disp, dcnt, cmds, cnt):
        b.free()
    return np.ascontiguousarray(recs).view(np.uint32).reshape(-1, 3), nd, H.cmds_u32(out), n


def run_oracle(sc, v, tests, opyr=None, dcap=1 << 21, cap=1 << 21, transforms=None, objects=None, object_hiz=False):
    """The object pass (PreClusterDrawCull.comp.glsl:13-47: frustum + LOD; with object_hiz also the Hi-Z test
    against opyr), the expansion, then the cluster tests (Hi-Z against opyr when given)."""
    objects = sc.objects if objects is None else objects
    transforms = sc.transforms if transforms is None else transforms
    scene = (objects, transforms, sc.surfaces, sc.lods)
    vb = H.view_bytes(v)
    recs, nd = orc.cluster_expand(scene, vb, opyr if object_hiz else None, capacity=dcap)
    cmds, n = orc.cluster_cull(recs, objects, transforms, sc.clusters, vb, opyr, tests, capacity=cap)
    return recs, nd, cmds, n


def check(got, exp):
    g_recs, g_nd, g_cmds, g_n = got
    e_recs, e_nd, e_cmds, e_n = exp
    assert (g_nd, g_n) == (e_nd, e_n)
    np.testing.assert_array_equal(g_recs, e_recs)
    np.testing.assert_array_equal(g_cmds, e_cmds)


@pytest.mark.parametrize("tests", [0, A.CLUSTER_FRUSTUM, A.CLUSTER_CONE, A.CLUSTER_FRUSTUM | A.CLUSTER_CONE, ALL])
@pytest.mark.parametrize("with_pyr,object_hiz", [(False, False), (True, False), (True, True)])
def test_cluster_cull_bit_exact(ctx, cl_scene, pyramid, tests, with_pyr, object_hiz):
    h, sc = cl_scene
    v = H.view(pyr=(1024, 512))
    v.lodTarget = 2e-5  # finer LODs: more clusters per object
    p, opyr = pyramid
    got = run_gpu(ctx, h, v, tests, pyr=p if with_pyr else 0, flags=A.CLUSTER_OBJECT_HIZ if object_hiz else 0)
    exp = run_oracle(sc, v, tests, opyr if with_pyr else None, object_hiz=object_hiz)
    assert exp[1] > 10000
    check(got, exp)


def test_cluster_object_hiz_flag_checked(ctx, cl_scene):
    h, sc = cl_scene
    b = ctx.alloc(64)
    with pytest.raises(bz.BzError):  # the object Hi-Z test needs a pyramid
        ctx.cull_clusters(h, H.view(), 0, b, b, b, b, dispatch_capacity=1, capacity=1, flags=A.CLUSTER_OBJECT_HIZ)
    with pytest.raises(bz.BzError):  # unknown flag bits
        ctx.cull_clusters(h, H.view(), 0, b, b, b, b, dispatch_capacity=1, capacity=1, flags=0x10)
    b.free()


def test_cluster_capacities(ctx, cl_scene, pyramid):
    h, sc = cl_scene
    v = H.view(pyr=(1024, 512), yaw=0.8)
    v.lodTarget = 2e-5
    p, opyr = pyramid
    full = run_oracle(sc, v, ALL, opyr)
    assert full[1] > 20000
    for dcap, cap in [(full[1] // 2, 1 << 21), (1 << 21, 100), (1000, 10), (0, 0), (1, 1)]:
        check(run_gpu(ctx, h, v, ALL, pyr=p, dcap=dcap, cap=cap), run_oracle(sc, v, ALL, opyr, dcap=dcap, cap=cap))


def test_cluster_default_lod_target_and_poses(ctx, cl_scene):
    h, sc = cl_scene
    for pose in [dict(), dict(pos=(100.0, 20.0, -50.0), yaw=2.0), dict(fov=50.0, w=512, h=512, pitch=0.7)]:
        v = H.view(**pose)
        check(run_gpu(ctx, h, v, A.CLUSTER_FRUSTUM | A.CLUSTER_CONE),
              run_oracle(sc, v, A.CLUSTER_FRUSTUM | A.CLUSTER_CONE))


def test_cluster_transform_updates(ctx):
    """1% of the transforms get seeded rotation updates each frame; every frame matches the oracle."""
    sc = bz.synth_scene(H.SEED_C4 + 1, 30000, 128, 8, clusters=(16, 128))
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    ctx.scene_set_clusters(h, sc.clusters)
    v = H.view()
    v.lodTarget = 1e-4
    xf = sc.transforms.copy()
    for frame in range(3):
        ids, new = bz.synth_transform_updates(5, frame, xf, len(xf) // 100)
        ctx.update_transforms(h, ids, new)
        xf[ids] = new
        check(run_gpu(ctx, h, v, A.CLUSTER_FRUSTUM | A.CLUSTER_CONE),
              run_oracle(sc, v, A.CLUSTER_FRUSTUM | A.CLUSTER_CONE, transforms=xf))
    ctx.scene_destroy(h)


def test_cluster_shared_transforms(ctx):
    sc = bz.synth_scene(H.SEED_C4 + 2, 20000, 128, 8, clusters=(16, 128))
    objs = sc.objects.copy()
    objs["transformId"] = np.random.default_rng(2).integers(0, 5000, len(objs), dtype=np.uint32)
    xf = sc.transforms[:5000].copy()
    h = ctx.scene_create(objs, xf, sc.surfaces, sc.lods)
    ctx.scene_set_clusters(h, sc.clusters)
    v = H.view()
    v.lodTarget = 5e-5
    check(run_gpu(ctx, h, v, A.CLUSTER_FRUSTUM | A.CLUSTER_CONE),
          run_oracle(sc, v, A.CLUSTER_FRUSTUM | A.CLUSTER_CONE, transforms=xf, objects=objs))
    ctx.scene_destroy(h)


def test_cluster_requires_table(ctx):
    sc = H.c0_scene()
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    b = ctx.alloc(64)
    with pytest.raises(bz.BzError):
        ctx.cull_clusters(h, H.view(), 0, b, b, b, b, dispatch_capacity=1, capacity=1)
    ctx.scene_destroy(h)
    b.free()
