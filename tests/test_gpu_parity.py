"""HIP path vs the CPU oracle, bit-exact, through the C ABI (libbzcull.so). Needs a gfx950 GPU."""
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


def gpu_cull(ctx, scene_h, v, pass_, n, pyramid=0, flags=0, begin=0, count=0, vis_in=None, vis_out=None,
             capacity=None):
    cap = (count or n) if capacity is None else capacity
    cmds = ctx.alloc(max(cap, 1) * 24)
    cnt = ctx.alloc(4)
    cnt.fill_u32(0xDEADBEEF, 1)
    ctx.cull(scene_h, v, pass_, cmds, cnt, capacity=cap, pyramid=pyramid, flags=flags, begin=begin, n=count,
             vis_in=vis_in, vis_out=vis_out)
    ctx.check_kernels()
    out, total = ctx.read_commands(cmds, cnt, capacity=cap)
    cmds.free()
    cnt.free()
    return H.cmds_u32(out), total


def upload_pyramid(ctx, depth, base_w, base_h, mips):
    h, w = depth.shape
    buf = ctx.alloc(depth.nbytes)
    buf.upload(depth)
    p = ctx.pyramid_create(base_w, base_h, mips)
    ctx.pyramid_build(p, buf.ptr, w, h)
    ctx.check_kernels()
    return p, buf


def test_plain_c0_bit_exact(ctx):
    sc = H.c0_scene()
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    v = H.view()
    got, n = gpu_cull(ctx, h, v, A.PASS_PLAIN, len(sc.objects))
    exp, ne = orc.cull(H.as_tuple(sc), H.view_bytes(v), A.PASS_PLAIN)
    assert n == ne and n > 0
    np.testing.assert_array_equal(got, exp)
    ctx.scene_destroy(h)


@pytest.mark.parametrize("dims", [(1920, 1080, 1024, 512, 11), (1920, 1080, 1024, 1024, 10), (512, 512, 256, 256, 8),
                                  (2048, 1024, 1024, 512, 10), (1920, 1080, 960, 540, 10), (37, 23, 16, 8, 5),
                                  (4096, 2048, 2048, 1024, 12), (2048, 64, 1024, 32, 11), (64, 4096, 32, 2048, 12),
                                  (3000, 100, 2048, 64, 12)])
def test_pyramid_bit_exact(ctx, dims):
    w, h, bw, bh, mips = dims
    depth = H.c1_depth(w=w, h=h, occluders=256)
    p, buf = upload_pyramid(ctx, depth, bw, bh, mips)
    got, _ = ctx.pyramid_read(p, mips)
    exp = orc.pyramid_build(depth, bw, bh, mips)
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
    ctx.pyramid_destroy(p)
    buf.free()


def test_pyramid_rebuild_twice(ctx):
    """Self-resetting tail counter: a second build on the same pyramid gives the same result."""
    depth = H.c1_depth()
    p, buf = upload_pyramid(ctx, depth, 1024, 512, 11)
    a, _ = ctx.pyramid_read(p, 11)
    ctx.pyramid_build(p, buf.ptr, 1920, 1080)
    ctx.check_kernels()
    b, _ = ctx.pyramid_read(p, 11)
    np.testing.assert_array_equal(a, b)
    ctx.pyramid_destroy(p)


def test_device_depth_matches_host(ctx):
    v = H.view()
    for (w, h, occ, seed) in [(1920, 1080, 512, 1), (512, 512, 64, 2), (1000, 777, 1024, 3)]:
        host = bz.synth_depth(seed, w, h, occ, v.proj5, 0.1)
        buf = ctx.alloc(host.nbytes)
        ctx.synth_depth_device(seed, w, h, occ, v.proj5, 0.1, buf.ptr)
        ctx.sync()
        dev = buf.download(np.float32, w * h).reshape(h, w)
        np.testing.assert_array_equal(dev.view(np.uint32), host.view(np.uint32))
        buf.free()


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 262144 + 17])
def test_two_pass_bit_exact(ctx, n):
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    depth = H.c1_depth()
    p, buf = upload_pyramid(ctx, depth, 1024, 512, 11)
    texels = orc.pyramid_build(depth, 1024, 512, 11)
    opyr = orc.pyramid(texels, 1024, 512, 11)
    v = H.view(pyr=(1024, 512))
    vb = H.view_bytes(v)
    vis0 = bz.synth_visibility(H.SEED_C1, n, 0.5)
    ctx.set_visibility(h, vis0)
    # early
    got_e, ne = gpu_cull(ctx, h, v, A.PASS_EARLY, n)
    exp_e, nee = orc.cull(scene, vb, A.PASS_EARLY, vis_in=vis0)
    assert ne == nee
    np.testing.assert_array_equal(got_e, exp_e)
    # late (in place)
    got_l, nl = gpu_cull(ctx, h, v, A.PASS_LATE, n, pyramid=p)
    vis_exp = vis0.copy()
    exp_l, nle = orc.cull(scene, vb, A.PASS_LATE, pyr=opyr, vis_in=vis0, vis_out=vis_exp)
    assert nl == nle
    np.testing.assert_array_equal(got_l, exp_l)
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vis_exp)
    ctx.scene_destroy(h)
    ctx.pyramid_destroy(p)
    buf.free()


@pytest.mark.parametrize("pass_", [A.PASS_TEMPORAL, A.PASS_ONPC])
def test_other_passes_bit_exact(ctx, pass_):
    n = 100000
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    depth = H.c1_depth()
    p, buf = upload_pyramid(ctx, depth, 1024, 512, 11)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    v = H.view(pyr=(1024, 512))
    got, ng = gpu_cull(ctx, h, v, pass_, n, pyramid=p)
    exp, ne = orc.cull(scene, H.view_bytes(v), pass_, pyr=opyr)
    assert ng == ne
    np.testing.assert_array_equal(got, exp)
    ctx.scene_destroy(h)
    ctx.pyramid_destroy(p)


def test_early_with_previous_pyramid(ctx):
    n = 50000
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    depth = H.c1_depth(seed=77)
    p, buf = upload_pyramid(ctx, depth, 1024, 512, 11)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    v = H.view(pyr=(1024, 512))
    vis0 = bz.synth_visibility(5, n, 0.7)
    ctx.set_visibility(h, vis0)
    got, ng = gpu_cull(ctx, h, v, A.PASS_EARLY, n, pyramid=p, flags=A.CULL_EARLY_HIZ)
    exp, ne = orc.cull(scene, H.view_bytes(v), A.PASS_EARLY, pyr=opyr, flags=A.CULL_EARLY_HIZ, vis_in=vis0)
    assert ng == ne
    np.testing.assert_array_equal(got, exp)
    ctx.scene_destroy(h)
    ctx.pyramid_destroy(p)


def test_capacity_and_ranges(ctx):
    n = 20000
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    v = H.view()
    full, nf = orc.cull(scene, H.view_bytes(v), A.PASS_PLAIN)
    # capacity smaller than the count: exact count, first `capacity` commands
    got, ng = gpu_cull(ctx, h, v, A.PASS_PLAIN, n, capacity=nf // 3)
    assert ng == nf
    np.testing.assert_array_equal(got, full[:nf // 3])
    # zero capacity
    got, ng = gpu_cull(ctx, h, v, A.PASS_PLAIN, n, capacity=0)
    assert ng == nf and len(got) == 0
    # object sub-range
    got, ng = gpu_cull(ctx, h, v, A.PASS_PLAIN, n, begin=1234, count=5000)
    exp, ne = orc.cull(scene, H.view_bytes(v), A.PASS_PLAIN, begin=1234, count=5000)
    assert ng == ne
    np.testing.assert_array_equal(got, exp)
    ctx.scene_destroy(h)


def test_shared_transforms_and_out_of_place_visibility(ctx):
    """transformId != objectId (indirect transform stream) and separate visibility in/out buffers."""
    n = 30000
    sc = H.c1_scene(n)
    rng = np.random.default_rng(3)
    objects = sc.objects.copy()
    objects["transformId"] = rng.integers(0, n // 2, n, dtype=np.uint32)
    scene = (objects, sc.transforms[: n // 2].copy(), sc.surfaces, sc.lods)
    h = ctx.scene_create(*scene)
    depth = H.c1_depth()
    p, buf = upload_pyramid(ctx, depth, 1024, 512, 11)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    v = H.view(pyr=(1024, 512))
    vis0 = bz.synth_visibility(9, n, 0.5)
    vin = ctx.alloc(n * 4)
    vin.upload(vis0)
    vout = ctx.alloc(n * 4)
    vout.fill_u32(7)
    got, ng = gpu_cull(ctx, h, v, A.PASS_LATE, n, pyramid=p, vis_in=vin.ptr, vis_out=vout.ptr)
    vexp = np.zeros(n, np.uint32)
    exp, ne = orc.cull(scene, H.view_bytes(v), A.PASS_LATE, pyr=opyr, vis_in=vis0, vis_out=vexp)
    assert ng == ne
    np.testing.assert_array_equal(got, exp)
    np.testing.assert_array_equal(vout.download(np.uint32, n), vexp)
    np.testing.assert_array_equal(vin.download(np.uint32, n), vis0)  # input untouched
    ctx.scene_destroy(h)


@pytest.mark.parametrize("tuning", [(256, 2), (256, 4), (512, 2), (512, 4)])
def test_tunings_identical(ctx, tuning):
    n = 70000
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    v = H.view()
    old = ctx.get_tuning(0)
    ctx.set_tuning(0, *tuning)
    try:
        got, ng = gpu_cull(ctx, h, v, A.PASS_PLAIN, n)
    finally:
        ctx.set_tuning(0, old.block, old.items)
    exp, ne = orc.cull(scene, H.view_bytes(v), A.PASS_PLAIN)
    assert ng == ne
    np.testing.assert_array_equal(got, exp)
    ctx.scene_destroy(h)


def test_transform_update_then_cull(ctx):
    n = 40000
    sc = H.c1_scene(n)
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    v = H.view()
    xf = sc.transforms.copy()
    for frame in range(3):
        ids, new = bz.synth_transform_updates(11, frame, xf, n // 100)
        ctx.update_transforms(h, ids, new)
        xf[ids] = new
        got, ng = gpu_cull(ctx, h, v, A.PASS_PLAIN, n)
        exp, ne = orc.cull((sc.objects, xf, sc.surfaces, sc.lods), H.view_bytes(v), A.PASS_PLAIN)
        assert ng == ne
        np.testing.assert_array_equal(got, exp)
    ctx.scene_destroy(h)


def test_determinism_repeat(ctx):
    n = 300000
    sc = H.c1_scene(n)
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    v = H.view()
    first, n0 = gpu_cull(ctx, h, v, A.PASS_PLAIN, n)
    for _ in range(5):
        again, n1 = gpu_cull(ctx, h, v, A.PASS_PLAIN, n)
        assert n1 == n0
        np.testing.assert_array_equal(again, first)
    ctx.scene_destroy(h)


@pytest.mark.parametrize("n_surfaces", [65536, 65537])
def test_surface_id_width(ctx, n_surfaces):
    """Surface ids are kept 16-bit when every id fits (up to 65 536 surfaces) and 32-bit otherwise; both give the
    oracle's commands, including for the largest ids."""
    n = 150000
    sc = bz.synth_scene(H.SEED_C1 + n_surfaces, n, n_surfaces, 2)
    objects = sc.objects.copy()
    objects["surfaceId"][:1000] = n_surfaces - 1 - np.arange(1000, dtype=np.uint32)  # the top of the id range
    scene = (objects, sc.transforms, sc.surfaces, sc.lods)
    h = ctx.scene_create(*scene)
    v = H.view()
    v.lodTarget = 1e-4
    got, ng = gpu_cull(ctx, h, v, A.PASS_PLAIN, n)
    exp, ne = orc.cull(scene, H.view_bytes(v), A.PASS_PLAIN)
    assert ng == ne and ne > 1000
    np.testing.assert_array_equal(got, exp)
    ctx.scene_destroy(h)
