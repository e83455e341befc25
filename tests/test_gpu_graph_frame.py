"""A whole frame (pyramid build, EARLY and LATE with the frame cache) captured from a stream into a hipGraph and
replayed, as bench.py times it: every replay must give the commands, counts and visibility of the eager frame and of
the oracle. Between replays the program runs eager work of its own on the same stream and other buffers (a PLAIN pass
and another frame) and synchronizes, so a replay that depended on state left by the capture, or on device memory the
eager work rewrites, would show. Needs a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from blitzen_amd.hip import StreamGraph
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu
FC = A.CULL_FRAME_CACHE


def test_captured_frame_replays_equal_eager_and_oracle():
    n = 300000
    ctx = bz.Context(0)
    try:
        st = ctx.stream()
        sc = H.c1_scene(n)
        scene = H.as_tuple(sc)
        h = ctx.scene_create(*scene)
        h2 = ctx.scene_create(*scene)
        depth = H.c1_depth()
        dbuf = ctx.alloc(depth.nbytes)
        dbuf.upload(depth)
        pyr = ctx.pyramid_create(1024, 512, 11)
        pyr2 = ctx.pyramid_create(1024, 512, 11)
        v = H.view(pyr=(1024, 512))
        vis0 = bz.synth_visibility(H.SEED_C1, n, 0.5)
        vin, vout = ctx.alloc(n * 4), ctx.alloc(n * 4)
        vin.upload(vis0)
        ce, cl, cnt = ctx.alloc(n * 24), ctx.alloc(n * 24), ctx.alloc(8)
        cp, cntp = ctx.alloc(n * 24), ctx.alloc(4)

        class _P:
            def __init__(self, p):
                self.ptr = p

        c_e, c_l = _P(cnt.ptr), _P(cnt.ptr + 4)

        def frame(hs, p):
            ctx.pyramid_build(p, dbuf.ptr, 1920, 1080, st)
            ctx.cull(hs, v, A.PASS_EARLY, ce, c_e, capacity=n, flags=FC, vis_in=vin.ptr, stream=st)
            ctx.cull(hs, v, A.PASS_LATE, cl, c_l, capacity=n, pyramid=p, flags=FC, vis_in=vin.ptr, vis_out=vout.ptr,
                     stream=st)

        def outputs():
            ctx.sync(st)
            counts = cnt.download(np.uint32, 2, st)
            e = H.cmds_u32(ce.download(bz.INDIRECT_DRAW, int(counts[0]), st))
            l_ = H.cmds_u32(cl.download(bz.INDIRECT_DRAW, int(counts[1]), st))
            return counts.copy(), e, l_, vout.download(np.uint32, n, st)

        frame(h, pyr)  # eager, also sizes every scratch buffer before the capture
        ctx.check_kernels(st)
        eager = outputs()
        vb = H.view_bytes(v)
        opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
        e_exp = orc.cull(scene, vb, A.PASS_EARLY, vis_in=vis0)
        vexp = vis0.copy()
        l_exp = orc.cull(scene, vb, A.PASS_LATE, opyr, vis_in=vexp, vis_out=vexp)
        assert (int(eager[0][0]), int(eager[0][1])) == (e_exp[1], l_exp[1]) and e_exp[1] > 0 and l_exp[1] > 0
        np.testing.assert_array_equal(eager[1], e_exp[0])
        np.testing.assert_array_equal(eager[2], l_exp[0])
        np.testing.assert_array_equal(eager[3], vexp)

        g = StreamGraph(st, lambda: frame(h, pyr))
        for rep in range(4):
            # eager work of the program's own between replays: a PLAIN pass and a frame on the other scene copy
            # (whose outputs are then overwritten by the replay)
            ctx.cull(h2, v, A.PASS_PLAIN, cp, cntp, capacity=n, stream=st)
            frame(h2, pyr2)
            ce.fill_u32(0xFFFFFFFF, n * 6, st)
            cl.fill_u32(0xFFFFFFFF, n * 6, st)
            vout.fill_u32(7, n, st)
            ctx.check_kernels(st)
            g.launch(st)
            ctx.check_kernels(st)
            got = outputs()
            np.testing.assert_array_equal(got[0], eager[0], err_msg=f"replay {rep}")
            np.testing.assert_array_equal(got[1], eager[1], err_msg=f"replay {rep}")
            np.testing.assert_array_equal(got[2], eager[2], err_msg=f"replay {rep}")
            np.testing.assert_array_equal(got[3], eager[3], err_msg=f"replay {rep}")
        del g
    finally:
        ctx.close()
