"""Frame cache (BZ_CULL_FRAME_CACHE): an EARLY pass records the frustum survivors and the LATE pass of the same view
reads them. Every output (commands, counts, visibility) must equal the oracle's, and the cache must not be used when
the view, the range or the transforms changed in between. Needs a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu
FC = A.CULL_FRAME_CACHE


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


def run(ctx, h, v, pass_, n, pyr=0, flags=0, begin=0, count=0, vis_in=None, vis_out=None):
    cap = count or n
    cmds = ctx.alloc(max(cap, 1) * 24)
    cnt = ctx.alloc(4)
    cnt.fill_u32(0xDEADBEEF, 1)
    ctx.cull(h, v, pass_, cmds, cnt, capacity=cap, pyramid=pyr, flags=flags, begin=begin, n=count, vis_in=vis_in,
             vis_out=vis_out)
    ctx.check_kernels()
    out, total = ctx.read_commands(cmds, cnt, capacity=cap)
    cmds.free()
    cnt.free()
    return H.cmds_u32(out), total


def frame(ctx, h, scene, v, pyramid, vis0, begin=0, count=0, between=None, late_view=None, in_place=True):
    """EARLY (recording) then LATE (reading) vs the oracle's EARLY and LATE on the same inputs."""
    p, opyr = pyramid
    n = len(scene[0])
    ctx.set_visibility(h, vis0)
    e_got = run(ctx, h, v, A.PASS_EARLY, n, flags=FC, begin=begin, count=count)
    early_scene = scene
    if between:
        scene = between(scene)
    lv = late_view or v
    vin = vout = None
    if not in_place:
        vout_buf = ctx.alloc(n * 4)
        vout_buf.fill_u32(7)
        vout = vout_buf.ptr
    l_got = run(ctx, h, lv, A.PASS_LATE, n, pyr=p, flags=FC, begin=begin, count=count, vis_out=vout)
    vb, lvb = H.view_bytes(v), H.view_bytes(lv)
    e_exp = orc.cull(early_scene, vb, A.PASS_EARLY, begin=begin, count=count, vis_in=vis0)
    vexp = vis0.copy() if in_place else np.full(n, 7, np.uint32)
    l_exp = orc.cull(scene, lvb, A.PASS_LATE, opyr, begin=begin, count=count, vis_in=vis0, vis_out=vexp)
    assert e_got[1] == e_exp[1] and l_got[1] == l_exp[1]
    np.testing.assert_array_equal(e_got[0], e_exp[0])
    np.testing.assert_array_equal(l_got[0], l_exp[0])
    vis = ctx.get_visibility(h, n) if in_place else vout_buf.download(np.uint32, n)
    np.testing.assert_array_equal(vis, vexp)
    if not in_place:
        vout_buf.free()
    return l_got[1]


@pytest.mark.parametrize("n", [1, 63, 1000, 262144 + 17])
def test_frame_cache_two_pass(ctx, pyramid, n):
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    v = H.view(pyr=(1024, 512))
    frame(ctx, h, scene, v, pyramid, bz.synth_visibility(H.SEED_C1, n, 0.5))
    frame(ctx, h, scene, v, pyramid, bz.synth_visibility(H.SEED_C1 + 1, n, 0.2), in_place=False)
    ctx.scene_destroy(h)


@pytest.mark.parametrize("tuning", [(256, 2), (512, 4)])
def test_frame_cache_tunings_ranges_shared_transforms(ctx, pyramid, tuning):
    n = 90000
    sc = H.c1_scene(n)
    objs = sc.objects.copy()
    objs["transformId"] = np.random.default_rng(4).integers(0, n // 3, n, dtype=np.uint32)
    scene = (objs, sc.transforms[: n // 3].copy(), sc.surfaces, sc.lods)
    h = ctx.scene_create(*scene)
    old = ctx.get_tuning(0)
    ctx.set_tuning(0, *tuning)
    try:
        v = H.view(pyr=(1024, 512), yaw=0.4)
        frame(ctx, h, scene, v, pyramid, bz.synth_visibility(5, n, 0.5), begin=1234, count=50000)
    finally:
        ctx.set_tuning(0, old.block, old.items, old.grid_per_cu)
    ctx.scene_destroy(h)


def test_frame_cache_invalidated_by_transform_update(ctx, pyramid):
    n = 60000
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    v = H.view(pyr=(1024, 512))

    def update(sc_tuple):
        objects, xf, surfaces, lods = sc_tuple
        ids, new = bz.synth_transform_updates(3, 0, xf, n // 10)
        ctx.update_transforms(h, ids, new)
        xf2 = xf.copy()
        xf2[ids] = new
        return objects, xf2, surfaces, lods

    frame(ctx, h, scene, v, pyramid, bz.synth_visibility(6, n, 0.5), between=update)
    ctx.scene_destroy(h)


def test_frame_cache_not_used_for_another_view_or_range(ctx, pyramid):
    n = 60000
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    v = H.view(pyr=(1024, 512))
    frame(ctx, h, scene, v, pyramid, bz.synth_visibility(7, n, 0.5), late_view=H.view(pyr=(1024, 512), yaw=0.01))
    finer = H.view(pyr=(1024, 512))
    finer.lodTarget = v.lodTarget * 0.25  # same frustum, other LOD choices: the recorded LODs must not be used
    frame(ctx, h, scene, v, pyramid, bz.synth_visibility(11, n, 0.5), late_view=finer)
    # another range for the late pass: the full pass runs
    p, opyr = pyramid
    vis0 = bz.synth_visibility(8, n, 0.5)
    ctx.set_visibility(h, vis0)
    run(ctx, h, v, A.PASS_EARLY, n, flags=FC)
    got = run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC, begin=10, count=n - 20)
    vexp = vis0.copy()
    exp = orc.cull(scene, H.view_bytes(v), A.PASS_LATE, opyr, begin=10, count=n - 20, vis_in=vis0, vis_out=vexp)
    assert got[1] == exp[1]
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    # a late pass with the flag and no recording early pass before it on this scene also runs in full
    h2 = ctx.scene_create(*scene)
    ctx.set_visibility(h2, vis0)
    got = run(ctx, h2, v, A.PASS_LATE, n, pyr=p, flags=FC)
    vexp = vis0.copy()
    exp = orc.cull(scene, H.view_bytes(v), A.PASS_LATE, opyr, vis_in=vis0, vis_out=vexp)
    np.testing.assert_array_equal(got[0], exp[0])
    ctx.scene_destroy(h)
    ctx.scene_destroy(h2)


def late_exp(scene, v, opyr, vis):
    """The oracle's LATE pass reading and rewriting `vis` in place (LateDrawCull.comp.glsl:51,71)."""
    return orc.cull(scene, H.view_bytes(v), A.PASS_LATE, opyr, vis_in=vis, vis_out=vis)


def check(got, exp):
    assert got[1] == exp[1], (got[1], exp[1])
    np.testing.assert_array_equal(got[0], exp[0])


def test_frame_cache_reads_current_visibility(ctx, pyramid):
    """The cached LATE pass reads the visibility it is given, not what the recording EARLY pass read: another buffer,
    or the scene's buffer replaced in between, gives the oracle's result and still uses the records."""
    n = 50000
    scene = H.as_tuple(H.c1_scene(n))
    h = ctx.scene_create(*scene)
    v = H.view(pyr=(1024, 512))
    p, opyr = pyramid
    vis_a = bz.synth_visibility(9, n, 0.5)
    vis_b = bz.synth_visibility(10, n, 0.5)
    ctx.set_visibility(h, vis_a)
    run(ctx, h, v, A.PASS_EARLY, n, flags=FC)  # records, reading the scene's buffer (vis_a)
    other = ctx.alloc(n * 4)
    other.upload(vis_b)
    got = run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC, vis_in=other.ptr)  # reads vis_b in place
    vexp = vis_b.copy()
    check(got, late_exp(scene, v, opyr, vexp))
    np.testing.assert_array_equal(other.download(np.uint32, n), vexp)
    assert ctx.frame_cache_stats(h) == (1, 1)
    run(ctx, h, v, A.PASS_EARLY, n, flags=FC)
    ctx.set_visibility(h, vis_b)  # replaced between the passes
    got = run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC)
    vexp = vis_b.copy()
    check(got, late_exp(scene, v, opyr, vexp))
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    assert ctx.frame_cache_stats(h) == (2, 2)
    other.free()
    ctx.scene_destroy(h)


def test_frame_cache_late_after_late_in_place(ctx, pyramid):
    """EARLY(cache) -> LATE(cache, in place) -> LATE(cache): the second LATE reads the visibility the first one wrote,
    as the reference's LATE pass would, so it emits only what is newly visible relative to that."""
    n = 70000
    scene = H.as_tuple(H.c1_scene(n))
    h = ctx.scene_create(*scene)
    v = H.view(pyr=(1024, 512))
    p, opyr = pyramid
    vis0 = bz.synth_visibility(21, n, 0.5)
    ctx.set_visibility(h, vis0)
    check(run(ctx, h, v, A.PASS_EARLY, n, flags=FC),
          orc.cull(scene, H.view_bytes(v), A.PASS_EARLY, vis_in=vis0))
    vexp = vis0.copy()
    first = late_exp(scene, v, opyr, vexp)
    check(run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC), first)
    second = late_exp(scene, v, opyr, vexp)
    assert first[1] > 0 and second[1] == 0  # everything visible now was visible before the second pass
    check(run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC), second)
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    # an uncached LATE pass in place between them reads and writes the same buffer: the cached one after it agrees
    ctx.set_visibility(h, vis0)
    run(ctx, h, v, A.PASS_EARLY, n, flags=FC)
    vexp = vis0.copy()
    check(run(ctx, h, v, A.PASS_LATE, n, pyr=p), late_exp(scene, v, opyr, vexp))
    check(run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC), late_exp(scene, v, opyr, vexp))
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    assert ctx.frame_cache_stats(h) == (2, 3)
    ctx.scene_destroy(h)


def test_frame_cache_after_device_visibility_writes(ctx, pyramid):
    """Writes to the scene's visibility through its device pointer (taken before the EARLY pass) or by a batched LATE
    job between EARLY(cache) and LATE(cache) are seen by the cached LATE pass."""
    n = 65000
    scene = H.as_tuple(H.c1_scene(n))
    h = ctx.scene_create(*scene)
    v = H.view(pyr=(1024, 512))
    p, opyr = pyramid
    vptr = ctx.visibility_ptr(h)
    vis0 = bz.synth_visibility(31, n, 0.5)
    ctx.set_visibility(h, vis0)
    run(ctx, h, v, A.PASS_EARLY, n, flags=FC)
    # the caller's own writes: a range set to 0, another to 1, through the pointer handed out before the EARLY pass
    lib = bz.lib()
    ctx.check(lib.bz_memset_d32(ctx.handle, vptr + 4 * 1000, 0, 20000, None), "memset")
    ctx.check(lib.bz_memset_d32(ctx.handle, vptr + 4 * 30000, 1, 20000, None), "memset")
    vexp = vis0.copy()
    vexp[1000:21000] = 0
    vexp[30000:50000] = 1
    check(run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC), late_exp(scene, v, opyr, vexp))
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    # a batched LATE job rewriting the scene's visibility in place between the passes
    vis1 = bz.synth_visibility(32, n, 0.3)
    ctx.set_visibility(h, vis1)
    run(ctx, h, v, A.PASS_EARLY, n, flags=FC)
    bcmds, bcnt = ctx.alloc(n * 24), ctx.alloc(4)
    v2 = H.view(pyr=(1024, 512), yaw=0.3)
    ctx.cull_batch([ctx.view_job(h, v2, A.PASS_LATE, bcmds, bcnt, pyramid=p)])
    ctx.check_kernels()
    vexp = vis1.copy()
    late_exp(scene, v2, opyr, vexp)
    check(run(ctx, h, v, A.PASS_LATE, n, pyr=p, flags=FC), late_exp(scene, v, opyr, vexp))
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    assert ctx.frame_cache_stats(h) == (2, 2)
    bcmds.free()
    bcnt.free()
    ctx.scene_destroy(h)


@pytest.mark.parametrize("tuning", [(512, 2), (256, 2), (512, 4)])
def test_frame_cache_dense_survivors(ctx, pyramid, tuning):
    """Recorded tiles where most objects pass the frustum (more survivors than the cached late pass's first-round
    window, and more commands per tile than one emit round) next to sparse ones: outputs still equal the oracle's."""
    n = 120000
    sc = H.c1_scene(n)
    xf = sc.transforms.copy()
    rng = np.random.default_rng(12)
    dense = (np.arange(n) // 20000) % 2 == 0  # alternating dense and sparse runs of 20000 objects
    z = rng.uniform(20.0, 400.0, n).astype(np.float32)
    xf["pos"][dense, 2] = z[dense]
    xf["pos"][dense, 0] = (rng.uniform(-0.3, 0.3, n).astype(np.float32) * z)[dense]
    xf["pos"][dense, 1] = (rng.uniform(-0.2, 0.2, n).astype(np.float32) * z)[dense]
    scene = (sc.objects, xf, sc.surfaces, sc.lods)
    v = H.view(pyr=(1024, 512))
    _, n_plain = orc.cull(scene, H.view_bytes(v), A.PASS_PLAIN)
    assert n_plain > 0.4 * n  # the dense runs really are dense
    h = ctx.scene_create(*scene)
    old = ctx.get_tuning(0)
    ctx.set_tuning(0, *tuning)
    try:
        frame(ctx, h, scene, v, pyramid, bz.synth_visibility(13, n, 0.5))
        frame(ctx, h, scene, v, pyramid, bz.synth_visibility(14, n, 0.05), in_place=False)
    finally:
        ctx.set_tuning(0, old.block, old.items, old.grid_per_cu)
    ctx.scene_destroy(h)
