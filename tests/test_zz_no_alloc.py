"""bz_cull_noalloc: the cull pass without device allocation (fixed compaction scratch held by the library). Its outputs
must equal bz_cull's bit for bit, call after call (the fixed scratch keeps the previous call's contents), and sizes
beyond BZ_NOALLOC_MAX_OBJECTS are refused before anything runs. Needs a GPU; calls go through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from tests import helpers as H

pytestmark = pytest.mark.gpu

MAX_OBJECTS = 4194304


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
    yield p
    ctx.pyramid_destroy(p)
    buf.free()


def cull(ctx, noalloc, h, v, pass_, n, pyr=0, flags=0, vis_in=None, vis_out=None, begin=0, count=0):
    """One pass through bz_cull or bz_cull_noalloc; returns (commands as uint32 rows, exact count)."""
    cap = count or n
    cmds = ctx.alloc(max(cap, 1) * 24)
    cnt = ctx.alloc(4)
    cnt.fill_u32(0xDEADBEEF, 1)
    cmds.fill_u32(0xA5A5A5A5)
    a = A.CullArgs(scene=h, pyramid=pyr, pass_=pass_, flags=flags, view=C.pointer(v), cmds_dev=cmds.ptr,
                   count_dev=cnt.ptr, capacity=cap, object_begin=begin, object_count=count,
                   visibility_in_dev=vis_in, visibility_out_dev=vis_out)
    fn = bz.lib().bz_cull_noalloc if noalloc else bz.lib().bz_cull
    ctx.check(fn(ctx.handle, C.byref(a), None), "bz_cull_noalloc" if noalloc else "bz_cull")
    ctx.check_kernels()
    out, total = ctx.read_commands(cmds, cnt, capacity=cap)
    cmds.free()
    cnt.free()
    return H.cmds_u32(out), total


def both(ctx, h, v, pass_, n, vis0, pyr=0, flags=0, begin=0, count=0):
    """The same pass through both entry points from the same visibility; outputs and visibility must be equal."""
    res = []
    for noalloc in (False, True):
        ctx.set_visibility(h, vis0)
        r = cull(ctx, noalloc, h, v, pass_, n, pyr=pyr, flags=flags, begin=begin, count=count)
        res.append((r, ctx.get_visibility(h, n)))
    (a, va), (b, vb) = res
    assert a[1] == b[1]
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(va, vb)
    return a[1]


def test_noalloc_matches_bz_cull_two_calls(ctx, pyramid):
    """Two calls in a row, the second on a smaller scene: every pass equals bz_cull's output."""
    for n in (262144 + 77, 5000):
        sc = H.c1_scene(n)
        h = ctx.scene_create(*H.as_tuple(sc))
        v = H.view(pyr=(1024, 512))
        vis0 = bz.synth_visibility(H.SEED_C1 + n, n, 0.5)
        assert both(ctx, h, v, A.PASS_PLAIN, n, vis0) > 0
        assert both(ctx, h, v, A.PASS_EARLY, n, vis0) > 0
        assert both(ctx, h, v, A.PASS_LATE, n, vis0, pyr=pyramid) > 0
        both(ctx, h, v, A.PASS_ONPC, n, vis0, begin=n // 7, count=n // 2)
        both(ctx, h, v, A.PASS_TEMPORAL, n, vis0, pyr=pyramid)
        ctx.scene_destroy(h)


def test_noalloc_frame_cache(ctx, pyramid):
    """With the frame cache allocated by one bz_cull, the no-alloc EARLY records and its LATE reads the records:
    outputs equal bz_cull's; before the cache exists, the no-alloc call refuses with BZ_E_STATE."""
    n = 100000
    sc = H.c1_scene(n)
    h = ctx.scene_create(*H.as_tuple(sc))
    v = H.view(pyr=(1024, 512))
    vis0 = bz.synth_visibility(77, n, 0.5)
    cmds, cnt = ctx.alloc(n * 24), ctx.alloc(4)
    a = A.CullArgs(scene=h, pyramid=0, pass_=A.PASS_EARLY, flags=A.CULL_FRAME_CACHE, view=C.pointer(v),
                   cmds_dev=cmds.ptr, count_dev=cnt.ptr, capacity=n)
    assert bz.lib().bz_cull_noalloc(ctx.handle, C.byref(a), None) == -9  # BZ_E_STATE
    cmds.free()
    cnt.free()
    got = {}
    for noalloc in (False, True):
        ctx.set_visibility(h, vis0)
        e = cull(ctx, noalloc, h, v, A.PASS_EARLY, n, flags=A.CULL_FRAME_CACHE)
        late = cull(ctx, noalloc, h, v, A.PASS_LATE, n, pyr=pyramid, flags=A.CULL_FRAME_CACHE)
        got[noalloc] = (e, late, ctx.get_visibility(h, n))
    for k in range(2):
        assert got[False][k][1] == got[True][k][1]
        np.testing.assert_array_equal(got[False][k][0], got[True][k][0])
    np.testing.assert_array_equal(got[False][2], got[True][2])
    rec, reused = ctx.frame_cache_stats(h)
    assert rec == 2 and reused == 2
    ctx.scene_destroy(h)


def test_noalloc_refuses_oversize(ctx):
    """More than BZ_NOALLOC_MAX_OBJECTS objects in one call: BZ_E_INVALID with a message, before any launch."""
    n = MAX_OBJECTS + 1
    sc = H.c1_scene(n)
    h = ctx.scene_create(*H.as_tuple(sc))
    v = H.view()
    cnt = ctx.alloc(4)
    cnt.fill_u32(0xDEADBEEF, 1)
    a = A.CullArgs(scene=h, pyramid=0, pass_=A.PASS_PLAIN, flags=0, view=C.pointer(v), cmds_dev=None,
                   count_dev=cnt.ptr, capacity=0)
    assert bz.lib().bz_cull_noalloc(ctx.handle, C.byref(a), None) == -1
    assert b"BZ_NOALLOC_MAX_OBJECTS" in bz.lib().bz_last_error(ctx.handle)
    ctx.sync()
    assert int(cnt.download(np.uint32, 1)[0]) == 0xDEADBEEF  # nothing ran
    a.object_count = MAX_OBJECTS  # a sub-range at the limit is accepted
    ctx.check(bz.lib().bz_cull_noalloc(ctx.handle, C.byref(a), None), "bz_cull_noalloc")
    ctx.check_kernels()
    cnt.free()
    ctx.scene_destroy(h)
