"""The C ABI's argument checks: every rejected call returns a status (never aborts), records a message, and launches
nothing; the outputs it was given are left untouched. Needs a GPU (a context needs a device)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from tests import helpers as H

pytestmark = pytest.mark.gpu
SENTINEL = 0xA5A5A5A5


@pytest.fixture(scope="module")
def env():
    ctx = bz.Context(0)
    sc = H.c1_scene(5000)
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    p = ctx.pyramid_create(64, 32, 7)
    cmds, cnt = ctx.alloc(5000 * 24 + 64), ctx.alloc(16)
    yield ctx, h, p, cmds, cnt
    ctx.close()


def call(ctx, **kw):
    """bz_cull with CullArgs built from kw; returns the status."""
    v = kw.pop("view", H.view(pyr=(64, 32)))
    a = A.CullArgs(view=C.pointer(v), **kw)
    return A.load().bz_cull(ctx.handle, C.byref(a), None)


@pytest.mark.parametrize("case", ["scene", "pass", "flags", "count_null", "cmds_null", "misaligned", "range",
                                  "pyramid", "late_no_pyramid", "early_hiz_no_pyramid"])
def test_cull_rejects(env, case):
    ctx, h, p, cmds, cnt = env
    cnt.fill_u32(SENTINEL, 4)
    cmds.fill_u32(SENTINEL, 64)
    ctx.sync()
    args = dict(scene=h, pyramid=0, pass_=A.PASS_PLAIN, flags=0, cmds_dev=cmds.ptr, count_dev=cnt.ptr, capacity=5000)
    if case == "scene":
        args["scene"] = 987654
    elif case == "pass":
        args["pass_"] = 7
    elif case == "flags":
        args["flags"] = 0x80
    elif case == "count_null":
        args["count_dev"] = None
    elif case == "cmds_null":
        args["cmds_dev"] = None
    elif case == "misaligned":
        args["cmds_dev"] = cmds.ptr + 4
    elif case == "range":
        args.update(object_begin=4000, object_count=1001)
    elif case == "pyramid":
        args.update(pass_=A.PASS_LATE, pyramid=123456)
    elif case == "late_no_pyramid":
        args["pass_"] = A.PASS_LATE
    elif case == "early_hiz_no_pyramid":
        args.update(pass_=A.PASS_EARLY, flags=A.CULL_EARLY_HIZ)
    assert call(ctx, **args) == A.E_INVALID
    assert A.load().bz_last_error(ctx.handle)  # a message was recorded
    ctx.check_kernels()
    assert (cnt.download(np.uint32, 4) == SENTINEL).all()
    assert (cmds.download(np.uint32, 64) == SENTINEL).all()


def test_valid_edge_calls(env):
    ctx, h, p, cmds, cnt = env
    # an empty range writes a zero count
    cnt.fill_u32(SENTINEL, 1)
    assert call(ctx, scene=h, pass_=A.PASS_PLAIN, cmds_dev=cmds.ptr, count_dev=cnt.ptr, capacity=10,
                object_begin=5000, object_count=0) == A.OK
    ctx.check_kernels()
    assert int(cnt.download(np.uint32, 1)[0]) == 0
    # capacity 0 with no command buffer: only the exact count
    cnt.fill_u32(SENTINEL, 1)
    assert call(ctx, scene=h, pass_=A.PASS_PLAIN, cmds_dev=None, count_dev=cnt.ptr, capacity=0) == A.OK
    ctx.check_kernels()
    n_all = int(cnt.download(np.uint32, 1)[0])
    assert 0 < n_all < 5000
    # capacity smaller than the count: the count stays exact, only `capacity` commands are written
    cmds.fill_u32(SENTINEL)
    assert call(ctx, scene=h, pass_=A.PASS_PLAIN, cmds_dev=cmds.ptr, count_dev=cnt.ptr, capacity=7) == A.OK
    ctx.check_kernels()
    assert int(cnt.download(np.uint32, 1)[0]) == n_all
    words = cmds.download(np.uint32, cmds.nbytes // 4)
    assert (words[7 * 6:] == SENTINEL).all() and not (words[:7 * 6] == SENTINEL).any()


def test_other_entry_points_reject(env):
    ctx, h, p, cmds, cnt = env
    lib = A.load()
    out = C.c_uint64()
    assert lib.bz_pyramid_create(ctx.handle, 0, 32, 5, C.byref(out)) == A.E_INVALID
    assert lib.bz_pyramid_create(ctx.handle, 64, 32, 17, C.byref(out)) == A.E_INVALID
    assert lib.bz_pyramid_destroy(ctx.handle, 55555) == A.E_INVALID
    assert lib.bz_scene_destroy(ctx.handle, 55555) == A.E_INVALID
    # more objects than one scene holds: rejected before any array is read
    one = H.c1_scene(1)
    sc_out = C.c_uint64()
    assert lib.bz_scene_create(ctx.handle, A.ptr(one.objects), 0x80000001, A.ptr(one.transforms), 1,
                               A.ptr(one.surfaces), len(one.surfaces), A.ptr(one.lods), len(one.lods),
                               C.byref(sc_out)) == A.E_INVALID
    assert sc_out.value == 0
    assert lib.bz_pyramid_build(ctx.handle, 55555, cmds.ptr, 128, 64, None) == A.E_INVALID
    ids = np.array([10, 6000], np.uint32)  # 6000 is past the transform table
    xf = np.zeros(2, A.MESH_TRANSFORM)
    assert lib.bz_scene_update_transforms(ctx.handle, h, A.ptr(ids), A.ptr(xf), 2, None) == A.E_INVALID
    # a batch with one bad job launches nothing
    v = H.view(pyr=(64, 32))
    good = A.ViewJob(scene=h, pyramid=p, pass_=A.PASS_PLAIN, object_begin=0, object_count=100, capacity=100, view=v,
                     cmds_dev=cmds.ptr, count_dev=cnt.ptr)
    bad = A.ViewJob(scene=987654, pyramid=p, pass_=A.PASS_PLAIN, object_begin=0, object_count=100, capacity=100,
                    view=v, cmds_dev=cmds.ptr, count_dev=cnt.ptr + 4)
    cnt.fill_u32(SENTINEL, 2)
    jobs = (A.ViewJob * 2)(good, bad)
    assert lib.bz_cull_batch(ctx.handle, jobs, 2, None) == A.E_INVALID
    ctx.check_kernels()
    assert (cnt.download(np.uint32, 2) == SENTINEL).all()
    # the context stays usable after rejected calls
    assert call(ctx, scene=h, pass_=A.PASS_PLAIN, cmds_dev=cmds.ptr, count_dev=cnt.ptr, capacity=5000) == A.OK
    ctx.check_kernels()
