"""Transform streaming through the caller-filled staging slots (bz_scene_transform_staging / _commit_transforms):
every cull after a commit equals the oracle's on the updated transforms, slots alternate, bad ids are reported.
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


def plain(ctx, h, v, n):
    cmds = ctx.alloc(n * 24)
    cnt = ctx.alloc(4)
    ctx.cull(h, v, A.PASS_PLAIN, cmds, cnt, capacity=n)
    ctx.check_kernels()
    out, total = ctx.read_commands(cmds, cnt, capacity=n)
    cmds.free()
    cnt.free()
    return H.cmds_u32(out), total


@pytest.mark.parametrize("shared", [False, True])
def test_staged_updates_match_oracle(ctx, shared):
    n = 50000
    sc = H.c1_scene(n)
    objects, xf = sc.objects.copy(), sc.transforms.copy()
    if shared:  # transformId != objectId: updates address the transform table
        objects["transformId"] = np.random.default_rng(5).integers(0, n // 2, n, dtype=np.uint32)
        xf = xf[: n // 2].copy()
    h = ctx.scene_create(objects, xf, sc.surfaces, sc.lods)
    v = H.view()
    seen = set()
    for frame in range(5):  # slots alternate; several commits in a row without waiting
        ids, new = bz.synth_transform_updates(21, frame, xf, len(xf) // 50)
        sid, sxf = ctx.transform_staging(h, len(ids) + 7)
        seen.add(sid.ctypes.data)
        sid[: len(ids)] = ids
        sxf[: len(ids)] = new
        ctx.commit_transforms(h, len(ids))
        xf[ids] = new
        got, ng = plain(ctx, h, v, n)
        exp, ne = orc.cull((objects, xf, sc.surfaces, sc.lods), H.view_bytes(v), A.PASS_PLAIN)
        assert ng == ne
        np.testing.assert_array_equal(got, H.cmds_u32(exp))
    assert len(seen) == 2  # two slots used in turn
    ctx.scene_destroy(h)


def test_staging_reopen_commit_zero_and_errors(ctx):
    n = 20000
    sc = H.c1_scene(n)
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    xf = sc.transforms.copy()
    v = H.view()
    with pytest.raises(A.BzError) as e:
        ctx.commit_transforms(h, 1)  # nothing handed out
    assert e.value.status == A.E_STATE
    a, _ = ctx.transform_staging(h, 100)
    b, _ = ctx.transform_staging(h, 100)  # not committed yet: the same slot again
    assert a.ctypes.data == b.ctypes.data
    ctx.commit_transforms(h, 0)
    ctx.check_kernels()
    # an out-of-range id is skipped and reported; the other updates of the commit are applied
    ids, new = bz.synth_transform_updates(22, 0, xf, 200)
    sid, sxf = ctx.transform_staging(h, 201)
    sid[:200] = ids
    sxf[:200] = new
    sid[200] = n + 5
    sxf[200] = new[0]
    ctx.commit_transforms(h, 201)
    with pytest.raises(A.BzError) as e:
        ctx.check_kernels()
    assert e.value.status == A.E_INVALID
    ctx.check_kernels()  # reported once
    xf[ids] = new
    got, ng = plain(ctx, h, v, n)
    exp, ne = orc.cull((sc.objects, xf, sc.surfaces, sc.lods), H.view_bytes(v), A.PASS_PLAIN)
    assert ng == ne
    np.testing.assert_array_equal(got, H.cmds_u32(exp))
    with pytest.raises(A.BzError):
        ctx.transform_staging(h, 0)
    ctx.scene_destroy(h)


def test_staged_update_invalidates_frame_cache(ctx):
    n = 40000
    sc = H.c1_scene(n)
    h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
    depth = H.c1_depth()
    dbuf = ctx.alloc(depth.nbytes)
    dbuf.upload(depth)
    p = ctx.pyramid_create(1024, 512, 11)
    ctx.pyramid_build(p, dbuf.ptr, 1920, 1080)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    v = H.view(pyr=(1024, 512))
    vis0 = bz.synth_visibility(23, n, 0.5)
    ctx.set_visibility(h, vis0)
    cmds, cnt = ctx.alloc(n * 24), ctx.alloc(4)
    ctx.cull(h, v, A.PASS_EARLY, cmds, cnt, capacity=n, flags=A.CULL_FRAME_CACHE)
    xf = sc.transforms.copy()
    ids, new = bz.synth_transform_updates(24, 0, xf, n // 10)
    sid, sxf = ctx.transform_staging(h, len(ids))
    sid[:] = ids
    sxf[:] = new
    ctx.commit_transforms(h, len(ids))
    xf[ids] = new
    ctx.cull(h, v, A.PASS_LATE, cmds, cnt, capacity=n, pyramid=p, flags=A.CULL_FRAME_CACHE)
    ctx.check_kernels()
    got, ng = ctx.read_commands(cmds, cnt, capacity=n)
    vexp = vis0.copy()
    exp, ne = orc.cull((sc.objects, xf, sc.surfaces, sc.lods), H.view_bytes(v), A.PASS_LATE, opyr, vis_in=vis0,
                       vis_out=vexp)
    assert ng == ne
    np.testing.assert_array_equal(H.cmds_u32(got), H.cmds_u32(exp))
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    for b in (cmds, cnt, dbuf):
        b.free()
    ctx.pyramid_destroy(p)
    ctx.scene_destroy(h)


@pytest.mark.parametrize("staged", [False, True])
def test_duplicate_ids_last_occurrence_wins(ctx, staged):
    """Ids named several times in one update, in no particular order: every transform ends up as its last occurrence
    (the reference copies updates in call order, vulkanDraw.cpp:816-820), never a mix of two records, on both the
    host-array path and the staged path; the cull after it equals the oracle's on those transforms."""
    n = 40000
    sc = H.c1_scene(n)
    xf = sc.transforms.copy()
    h = ctx.scene_create(sc.objects, xf, sc.surfaces, sc.lods)
    rng = np.random.default_rng(41 + staged)
    for frame in range(3):
        ids = rng.integers(0, 300, 5000).astype(np.uint32)  # ~17 updates per id among 300 ids
        ids[-50:] = rng.integers(0, n, 50)                  # and some unique ones
        new = np.zeros(len(ids), dtype=xf.dtype)
        new["pos"] = rng.uniform(-60, 60, (len(ids), 3)).astype(np.float32)
        new["pos"][:, 2] = rng.uniform(1, 120, len(ids))
        new["scale"] = rng.uniform(0.5, 2, len(ids))
        q = rng.normal(size=(len(ids), 4)).astype(np.float32)
        new["orientation"] = q / np.linalg.norm(q, axis=1, keepdims=True)
        if staged:
            sid, sxf = ctx.transform_staging(h, len(ids))
            sid[:] = ids
            sxf[:] = new
            ctx.commit_transforms(h, len(ids))
        else:
            ctx.update_transforms(h, ids, new)
        ctx.check_kernels()
        for i, t in zip(ids, new):  # in order: the last occurrence stays
            xf[i] = t
        np.testing.assert_array_equal(ctx.get_transforms(h, 0, n).view(np.uint32), xf.view(np.uint32))
    v = H.view()
    got, ng = plain(ctx, h, v, n)
    exp, ne = orc.cull((sc.objects, xf, sc.surfaces, sc.lods), H.view_bytes(v), A.PASS_PLAIN)
    assert ng == ne
    np.testing.assert_array_equal(got, H.cmds_u32(exp))
    ctx.scene_destroy(h)
