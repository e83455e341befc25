"""Bit-exact parity where the rounding decides: objects whose bounding spheres touch the frustum's side planes
(tests.helpers.boundary_scene). On these scenes the oracle and the same expressions without fused multiply-adds
decide a few percent of the objects differently (tests/test_oracle_math.py), so a kernel that fused at other points
than the oracle, or rounded any step of the transform or the plane test differently, fails here. Covers the
single-view decide (PLAIN), the frame-cache pair (EARLY recording the survivors' spheres, the cached LATE projecting
them onto the Hi-Z pyramid) and the batched decide (TEMPORAL). The second half does the same for the Hi-Z level:
spheres whose projected box is within a rounding of a power of two in texels (tests.helpers.hiz_level_scene), tested
against a pyramid whose even levels pass and odd levels occlude, so every level choice shows in the output. Needs a
GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu
N = 65536
VIEWS = [(0.3, 0.1, 7), (-1.1, -0.4, 8)]


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


def cull(ctx, h, v, pass_, pyr=0, flags=0):
    cmds, cnt = ctx.alloc(N * 24), ctx.alloc(4)
    cnt.fill_u32(0xDEADBEEF, 1)
    ctx.cull(h, v, pass_, cmds, cnt, capacity=N, pyramid=pyr, flags=flags)
    ctx.check_kernels()
    out, total = ctx.read_commands(cmds, cnt, capacity=N)
    cmds.free()
    cnt.free()
    return H.cmds_u32(out), total


@pytest.mark.parametrize("yaw,pitch,seed", VIEWS)
def test_plain_on_the_side_planes(ctx, yaw, pitch, seed):
    v = H.view(yaw=yaw, pitch=pitch)
    scene = H.boundary_scene(N, v, seed)
    h = ctx.scene_create(*scene)
    got = cull(ctx, h, v, A.PASS_PLAIN)
    exp = orc.cull(scene, H.view_bytes(v), A.PASS_PLAIN, capacity=N)
    assert got[1] == exp[1] and 0.25 * N < got[1] < 0.85 * N
    np.testing.assert_array_equal(got[0], exp[0])
    ctx.scene_destroy(h)


@pytest.mark.parametrize("yaw,pitch,seed", VIEWS)
def test_frame_cache_on_the_side_planes(ctx, pyramid, yaw, pitch, seed):
    p, opyr = pyramid
    v = H.view(yaw=yaw, pitch=pitch, pyr=(1024, 512))
    scene = H.boundary_scene(N, v, seed)
    h = ctx.scene_create(*scene)
    vis0 = bz.synth_visibility(seed, N, 0.5)
    ctx.set_visibility(h, vis0)
    vb = H.view_bytes(v)
    e_got = cull(ctx, h, v, A.PASS_EARLY, flags=A.CULL_FRAME_CACHE)
    l_got = cull(ctx, h, v, A.PASS_LATE, pyr=p, flags=A.CULL_FRAME_CACHE)
    assert ctx.frame_cache_stats(h)[1] >= 1  # the LATE pass read the recorded spheres
    e_exp = orc.cull(scene, vb, A.PASS_EARLY, vis_in=vis0, capacity=N)
    vexp = vis0.copy()
    l_exp = orc.cull(scene, vb, A.PASS_LATE, opyr, vis_in=vis0, vis_out=vexp, capacity=N)
    assert e_got[1] == e_exp[1] and l_got[1] == l_exp[1] and l_got[1] > 0
    np.testing.assert_array_equal(e_got[0], e_exp[0])
    np.testing.assert_array_equal(l_got[0], l_exp[0])
    np.testing.assert_array_equal(ctx.get_visibility(h, N), vexp)
    ctx.scene_destroy(h)


def test_batched_on_the_side_planes(ctx, pyramid):
    """Two jobs of one batch over the two halves of a boundary scene of 2 N objects, both with the first view."""
    p, opyr = pyramid
    yaw, pitch, seed = VIEWS[0]
    v = H.view(yaw=yaw, pitch=pitch, pyr=(1024, 512))
    scene = H.boundary_scene(2 * N, v, seed)
    h = ctx.scene_create(*scene)
    bufs, jobs = [], []
    for j in range(2):
        cmds, cnt = ctx.alloc(N * 24), ctx.alloc(4)
        bufs.append((cmds, cnt))
        jobs.append(ctx.view_job(h, v, A.PASS_TEMPORAL, cmds, cnt, capacity=N, pyramid=p, begin=j * N, n=N))
    ctx.cull_batch(jobs)
    ctx.check_kernels()
    for j, (cmds, cnt) in enumerate(bufs):
        out, total = ctx.read_commands(cmds, cnt, capacity=N)
        exp = orc.cull(scene, H.view_bytes(v), A.PASS_TEMPORAL, opyr, begin=j * N, count=N, capacity=N)
        assert total == exp[1] and total > 0
        np.testing.assert_array_equal(H.cmds_u32(out), exp[0])
        cmds.free()
        cnt.free()
    ctx.scene_destroy(h)


# ------------------------------------------------------------------------------ Hi-Z level boundaries
@pytest.fixture(scope="module")
def parity_pyramid(ctx):
    """A 1024 x 512, 11-level pyramid holding tests.helpers.parity_pyramid (even levels 0.0, odd levels 1.0), written
    into the pyramid's storage directly."""
    texels = H.parity_pyramid(1024, 512, 11)
    p = ctx.pyramid_create(1024, 512, 11)
    tex, n = C.c_void_p(), C.c_uint64()
    offs = (C.c_uint32 * 11)()
    ctx.check(bz.lib().bz_pyramid_storage(ctx.handle, p, C.byref(tex), C.byref(n), offs), "bz_pyramid_storage")
    m = min(n.value, texels.size)
    ctx.check(bz.lib().bz_memcpy_h2d(ctx.handle, tex.value, A.ptr(texels[:m]), m * 4, None), "bz_memcpy_h2d")
    got, _ = ctx.pyramid_read(p, 11)
    np.testing.assert_array_equal(got[:m], texels[:m])
    yield p, orc.pyramid(texels, 1024, 512, 11)
    ctx.pyramid_destroy(p)


@pytest.mark.parametrize("yaw,pitch,seed", VIEWS)
def test_temporal_at_hiz_level_boundaries(ctx, parity_pyramid, yaw, pitch, seed):
    p, opyr = parity_pyramid
    v = H.view(yaw=yaw, pitch=pitch, pyr=(1024, 512))
    scene = H.hiz_level_scene(N, v, seed)
    h = ctx.scene_create(*scene)
    got = cull(ctx, h, v, A.PASS_TEMPORAL, pyr=p)
    exp = orc.cull(scene, H.view_bytes(v), A.PASS_TEMPORAL, opyr, capacity=N)
    assert got[1] == exp[1] and 0.1 * N < got[1] < 0.9 * N
    np.testing.assert_array_equal(got[0], exp[0])
    ctx.scene_destroy(h)


@pytest.mark.parametrize("yaw,pitch,seed", VIEWS)
def test_frame_cache_at_hiz_level_boundaries(ctx, parity_pyramid, yaw, pitch, seed):
    """Nothing was visible last frame: EARLY (recording) emits nothing and the cached LATE emits every object whose
    recorded sphere projects to an even level."""
    p, opyr = parity_pyramid
    v = H.view(yaw=yaw, pitch=pitch, pyr=(1024, 512))
    scene = H.hiz_level_scene(N, v, seed)
    h = ctx.scene_create(*scene)
    vis0 = np.zeros(N, np.uint32)
    ctx.set_visibility(h, vis0)
    e_got = cull(ctx, h, v, A.PASS_EARLY, flags=A.CULL_FRAME_CACHE)
    l_got = cull(ctx, h, v, A.PASS_LATE, pyr=p, flags=A.CULL_FRAME_CACHE)
    assert e_got[1] == 0 and ctx.frame_cache_stats(h)[1] >= 1
    vexp = vis0.copy()
    l_exp = orc.cull(scene, H.view_bytes(v), A.PASS_LATE, opyr, vis_in=vis0, vis_out=vexp, capacity=N)
    assert l_got[1] == l_exp[1] and 0.1 * N < l_got[1] < 0.9 * N
    np.testing.assert_array_equal(l_got[0], l_exp[0])
    np.testing.assert_array_equal(ctx.get_visibility(h, N), vexp)
    ctx.scene_destroy(h)


def test_batched_at_hiz_level_boundaries(ctx, parity_pyramid):
    p, opyr = parity_pyramid
    yaw, pitch, seed = VIEWS[1]
    v = H.view(yaw=yaw, pitch=pitch, pyr=(1024, 512))
    scene = H.hiz_level_scene(2 * N, v, seed)
    h = ctx.scene_create(*scene)
    bufs, jobs = [], []
    for j in range(2):
        cmds, cnt = ctx.alloc(N * 24), ctx.alloc(4)
        bufs.append((cmds, cnt))
        jobs.append(ctx.view_job(h, v, A.PASS_TEMPORAL, cmds, cnt, capacity=N, pyramid=p, begin=j * N, n=N))
    ctx.cull_batch(jobs)
    ctx.check_kernels()
    for j, (cmds, cnt) in enumerate(bufs):
        out, total = ctx.read_commands(cmds, cnt, capacity=N)
        exp = orc.cull(scene, H.view_bytes(v), A.PASS_TEMPORAL, opyr, begin=j * N, count=N, capacity=N)
        assert total == exp[1] and total > 0
        np.testing.assert_array_equal(H.cmds_u32(out), exp[0])
        cmds.free()
        cnt.free()
    ctx.scene_destroy(h)
