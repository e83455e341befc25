"""LOD-instanced cull (the D3D12 instanced path) on the GPU vs the CPU oracle, bit-exact. Needs a GPU."""
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


def run(ctx, h, scene, v, per_lod, capacity=None):
    nl = len(scene[3])
    cap = nl if capacity is None else capacity
    ids = ctx.alloc(nl * per_lod * 4)
    ids.fill_u32(0xFFFFFFFF)
    counters = ctx.alloc(nl * 8)
    cmds = ctx.alloc(max(cap, 1) * 32)
    cnt = ctx.alloc(4)
    cnt.fill_u32(0xDEADBEEF, 1)
    ctx.cull_instanced(h, v, per_lod, ids, counters, cmds, cnt, capacity=cap)
    ctx.check_kernels()
    n = int(cnt.download(np.uint32, 1)[0])
    out = (ids.download(np.uint32, nl * per_lod).reshape(nl, per_lod),
           counters.download(np.uint32, nl * 2).reshape(nl, 2),
           cmds.download(np.uint32, min(n, cap) * 8).reshape(-1, 8), n)
    for b in (ids, counters, cmds, cnt):
        b.free()
    return out


def check(got, exp):
    for g, e in zip(got[:3], exp[:3]):
        np.testing.assert_array_equal(g, e)
    assert got[3] == exp[3]


@pytest.mark.parametrize("n,surfaces,lods_per,per_lod", [(16384, 64, 4, 100000), (300000, 256, 8, 2000),
                                                         (200000, 256, 8, 40), (50000, 600, 4, 500),
                                                         (60000, 2500, 4, 100), (1, 3, 2, 10)])
def test_instanced_bit_exact(ctx, n, surfaces, lods_per, per_lod):
    sc = bz.synth_scene(H.SEED_C0 + n, n, surfaces, lods_per)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    v = H.view()
    v.lodTarget = 1e-4
    check(run(ctx, h, scene, v, per_lod), orc.cull_instanced(scene, H.view_bytes(v), per_lod))
    check(run(ctx, h, scene, v, per_lod, capacity=2), orc.cull_instanced(scene, H.view_bytes(v), per_lod, capacity=2))
    ctx.scene_destroy(h)


def test_instanced_tunings_and_shared_transforms(ctx):
    n = 120000
    sc = H.c1_scene(n)
    objs = sc.objects.copy()
    objs["transformId"] = np.random.default_rng(8).integers(0, n // 4, n, dtype=np.uint32)
    scene = (objs, sc.transforms[: n // 4].copy(), sc.surfaces, sc.lods)
    h = ctx.scene_create(*scene)
    v = H.view(yaw=1.1)
    v.lodTarget = 2e-4
    old = ctx.get_tuning(0)
    try:
        for tun in [(256, 2), (512, 4)]:
            ctx.set_tuning(0, *tun)
            check(run(ctx, h, scene, v, 300), orc.cull_instanced(scene, H.view_bytes(v), 300))
    finally:
        ctx.set_tuning(0, old.block, old.items, old.grid_per_cu)
    ctx.scene_destroy(h)


def test_instanced_more_than_64_tiles_per_chunk(ctx):
    """A chunk of decided tiles is read in windows of 64 tiles: with 512-object tiles, 17 039 360 objects make
    33 280 tiles, 65 per chunk (the chunk count is capped at 512), so every chunk spans two windows."""
    n = 512 * 512 * 65
    sc = bz.synth_scene(H.SEED_C0 + 7, n, 64, 4)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    v = H.view()
    v.lodTarget = 1e-4
    old = ctx.get_tuning(0)
    try:
        ctx.set_tuning(0, 256, 2)
        check(run(ctx, h, scene, v, 64), orc.cull_instanced(scene, H.view_bytes(v), 64))
    finally:
        ctx.set_tuning(0, old.block, old.items, old.grid_per_cu)
    ctx.scene_destroy(h)


def test_instanced_scene_without_objects(ctx):
    """No objects: no commands, every counter {lod * per_lod, 0}, no list slot written."""
    sc = bz.synth_scene(H.SEED_C0, 16, 8, 3)
    scene = (sc.objects[:0].copy(), sc.transforms, sc.surfaces, sc.lods)
    h = ctx.scene_create(*scene)
    got = run(ctx, h, scene, H.view(), 5)
    check(got, orc.cull_instanced(scene, H.view_bytes(H.view()), 5))
    assert got[3] == 0 and (got[1][:, 1] == 0).all() and (got[0] == 0xFFFFFFFF).all()
    ctx.scene_destroy(h)
