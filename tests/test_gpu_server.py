"""Serving loop on the GPU: each request's commands equal the oracle's and are the same alone and in a mix of
concurrently batched requests; budget admission, deadlines, cancellation and statistics. Needs a GPU."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

SIZES = [(512, 512), (1024, 768), (1920, 1080), (800, 600)]


@pytest.fixture(scope="module")
def ctx():
    c = bz.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def served(ctx):
    scenes = []
    for k, n in enumerate([40000, 120000, 7]):
        sc = H.c1_scene(n, seed=H.SEED_C1 + 100 + k)
        scenes.append((ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods), sc))
    ctx.server_start(2 << 30, inflight=3, max_batch=8)
    yield scenes
    ctx.server_stop()
    for h, _ in scenes:
        ctx.scene_destroy(h)


def make_requests(scenes, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        h, sc = scenes[i % len(scenes)]
        w, hh = SIZES[int(rng.integers(len(SIZES)))]
        v = H.view(fov=float(rng.uniform(50, 90)), w=w, h=hh, pos=tuple(rng.uniform(-30, 30, 3)),
                   yaw=float(rng.uniform(-3, 3)), pitch=float(rng.uniform(-0.4, 0.4)))
        pass_ = A.PASS_PLAIN if i % 5 == 4 else A.PASS_TEMPORAL
        out.append((bz.Context.request(h, v, depth_seed=int(seed * 1000 + i), depth_w=w, depth_h=hh,
                                       occluders=int(rng.integers(64, 513)), pass_=pass_, user_tag=i), sc))
    return out


def oracle_for(req, sc):
    v = bz.CameraViewData()
    C.memmove(C.addressof(v), C.addressof(req.view), C.sizeof(v))
    w, h = req.depth_w, req.depth_h
    bw, bh, mips = bz.reference_pyramid_extent(w, h)
    v.pyramidWidth, v.pyramidHeight = float(bw), float(bh)
    pyr = None
    if req.pass_ == A.PASS_TEMPORAL:
        depth = bz.synth_depth(req.depth_seed, w, h, req.occluders, v.proj5, v.zNear)
        pyr = orc.pyramid(orc.pyramid_build(depth, bw, bh, mips), bw, bh, mips)
    return orc.cull(H.as_tuple(sc), H.view_bytes(v), req.pass_, pyr, threads=8)


def run_and_fetch(ctx, tickets, reqs):
    out = []
    for t, (req, sc) in zip(tickets, reqs):
        r = ctx.wait(int(t), timeout_s=60)
        assert r.status == A.OK, r.status
        assert r.user_tag == req.user_tag
        cmds, n = ctx.fetch(int(t), len(sc.objects))
        assert n == r.count
        out.append((H.cmds_u32(cmds), n, r))
    return out


def test_single_request_matches_oracle(ctx, served):
    for req, sc in make_requests(served, 4, 1):
        t = ctx.submit(req)
        (got, n, r), = run_and_fetch(ctx, [t], [(req, sc)])
        exp, ne = oracle_for(req, sc)
        assert n == ne
        np.testing.assert_array_equal(got, exp)
        assert r.arrival_ns == r.submit_ns <= r.admit_ns <= r.complete_ns


def test_same_alone_and_in_the_mix(ctx, served):
    reqs = make_requests(served, 40, 2)
    alone = []
    for req, sc in reqs:
        alone += run_and_fetch(ctx, [ctx.submit(req)], [(req, sc)])
    tickets = [ctx.submit(req) for req, _ in reqs]  # all at once: batched while others are in flight
    mix = run_and_fetch(ctx, tickets, reqs)
    assert max(r.batch_size for _, _, r in mix) > 1
    for (a, na, _), (b, nb, _) in zip(alone, mix):
        assert na == nb
        np.testing.assert_array_equal(a, b)
    for k in (0, 7, 13):  # and against the oracle
        exp, ne = oracle_for(*reqs[k])
        assert alone[k][1] == ne
        np.testing.assert_array_equal(alone[k][0], exp)


def test_scheduled_arrivals(ctx, served):
    reqs = make_requests(served, 30, 3)
    offsets = np.cumsum(np.random.default_rng(3).exponential(1e9 / 5000, len(reqs))).astype(np.uint64)
    tickets, statuses = ctx.submit_schedule([r for r, _ in reqs], offsets)
    assert (statuses == A.OK).all() and (tickets > 0).all()
    res = run_and_fetch(ctx, tickets, reqs)
    for k in (0, 29):
        exp, ne = oracle_for(*reqs[k])
        np.testing.assert_array_equal(res[k][0], exp)
    sub = [r.submit_ns for _, _, r in res]
    assert all(b >= a for a, b in zip(sub, sub[1:]))
    # arrivals are the scheduled times: offsets apart, never after the actual submission
    arr = np.array([r.arrival_ns for _, _, r in res], np.uint64)
    np.testing.assert_array_equal(arr - arr[0], offsets - offsets[0])
    assert all(r.arrival_ns <= r.submit_ns for _, _, r in res)


def test_deadline_cancel_release_stats(ctx, served):
    h, sc = served[1]
    before = ctx.server_stats()
    past = bz.Context.request(h, H.view(), 5, 512, 512, 64, deadline_ns=1)  # long past
    t = ctx.submit(past)
    r = ctx.wait(t, timeout_s=30)
    assert r.status == A.E_TIMEOUT
    ctx.release(t)
    many = [ctx.submit(bz.Context.request(h, H.view(yaw=0.1 * k), 100 + k, 1920, 1080, 512)) for k in range(20)]
    try:
        ctx.cancel(many[-1])
        cancelled = True
    except bz.BzError:
        cancelled = False  # it had already completed
    statuses = [ctx.wait(int(t), timeout_s=60).status for t in many]
    if cancelled:
        assert statuses[-1] == A.E_CANCELLED
    assert all(s == A.OK for s in statuses[:-1])
    for t in many:
        ctx.release(int(t))
    st = ctx.server_stats()
    assert st.submitted == before.submitted + 21
    assert st.budget_in_use == 0 and st.budget_peak > 0 and st.batches > before.batches


def test_budget_admission():
    """A request larger than the whole budget is rejected; requests that fit one at a time all complete."""
    c = bz.Context(0)
    try:
        sc = H.c1_scene(300000, seed=77)
        h = c.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
        big = H.c1_scene(2000000, seed=78)
        hb = c.scene_create(big.objects, big.transforms, big.surfaces, big.lods)
        # a 1920x1080 request over 300 000 objects needs ~21 MiB (depth, pyramid, commands): one fits at a time
        c.server_start(32 << 20, inflight=2, max_batch=4)
        with pytest.raises(bz.BzError) as ei:
            c.submit(bz.Context.request(hb, H.view(), 1, 1920, 1080, 64))  # 48 MB of commands alone
        assert ei.value.status == A.E_BUDGET
        ts = [c.submit(bz.Context.request(h, H.view(yaw=0.2 * k), 10 + k, 1920, 1080, 128)) for k in range(6)]
        res = []
        for t in ts:  # a client that releases each result as it arrives frees room for the next request
            res.append(c.wait(int(t), timeout_s=60))
            c.release(int(t))
        assert all(r.status == A.OK for r in res)
        assert max(r.batch_size for r in res) == 1  # admitted one at a time
        st = c.server_stats()
        assert st.budget_peak <= (32 << 20) and st.rejected == 1
        c.server_stop()
    finally:
        c.close()
