"""The give-up path of the single-pass pyramid build, forced from the host: with a wait limit of 0 sleeps the tail
workgroup gives up at the first tile hand-off word not yet stored (every kernel still exits normally). bz_check
must report it, and must clear the words that tiles stored after the give-up, so that the next build with the
default limit equals the oracle's pyramid. That build uses a depth image other than the failed builds' one: a
leftover word would carry the old image's tile value into it. Needs a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_give_up_is_reported_and_the_next_build_is_exact():
    ctx = bz.Context(0)
    try:
        old = H.c1_depth(seed=H.SEED_C1 + 11, occluders=64)
        new = H.c1_depth()
        buf = ctx.alloc(old.nbytes)
        p = ctx.pyramid_create(1024, 512, 11)
        ctx.set_wait_limit(0)
        gave_up = 0
        buf.upload(old)
        for _ in range(4):  # each build completes or gives up; a give-up is reported by the next check only
            ctx.pyramid_build(p, buf.ptr, 1920, 1080)
            try:
                ctx.check_kernels()
            except A.BzError as e:
                assert e.status == A.E_INTERNAL and "bounded wait" in str(e)
                gave_up += 1
        print(f"builds that gave up: {gave_up} of 4")
        ctx.set_wait_limit(A.DEFAULT_WAIT_LIMIT)
        ctx.check_kernels()  # the flag was cleared by the check that reported it
        buf.upload(new)
        ctx.pyramid_build(p, buf.ptr, 1920, 1080)
        ctx.check_kernels()
        got, _ = ctx.pyramid_read(p, 11)
        exp = orc.pyramid_build(new, 1024, 512, 11)
        np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
        ctx.pyramid_destroy(p)
    finally:
        ctx.close()
