"""Vulkan interop: memory exported as a dma-buf file descriptor, imported through bz_extmem_*, and used as the
depth input and the command / count outputs of a two-pass frame. No Vulkan driver is present, so the exporter here
is the HIP runtime itself (hipMemGetHandleForAddressRange on a device allocation, the same kind of descriptor a
Vulkan driver on amdgpu exports). Results are read back through the exporter's pointer and must equal the oracle
bit for bit. Needs a GPU."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu
MiB = 1 << 20


@pytest.fixture(scope="module")
def ctx():
    c = bz.Context(0)
    yield c
    c.close()


def export_dmabuf(buf) -> int:
    """A dma-buf descriptor for the whole allocation (the exporter's side of the hand-over)."""
    hip = C.CDLL("libamdhip64.so.7")
    fn = hip.hipMemGetHandleForAddressRange
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_ulonglong]
    fd = C.c_int(-1)
    e = fn(C.byref(fd), buf.ptr, buf.nbytes, 1, 0)  # hipMemRangeHandleTypeDmaBufFd
    if e != 0 or fd.value < 0:
        pytest.skip(f"dma-buf export of device memory is not available here (hipError {e})")
    return fd.value


def test_import_map_read_write(ctx):
    src = ctx.alloc(4 * MiB)
    pattern = (np.arange(MiB, dtype=np.uint64) * 2654435761 % (1 << 32)).astype(np.uint32)
    src.upload(pattern)
    fd = export_dmabuf(src)
    mem = ctx.extmem_import(fd, src.nbytes)
    os.close(fd)  # the library imported its own duplicate
    view = ctx.extmem_map(mem, 1 * MiB, 2 * MiB)
    ctx.sync()
    np.testing.assert_array_equal(view.download(np.uint32, MiB // 2), pattern[MiB // 4: 3 * MiB // 4])
    view.fill_u32(0xA5A5A5A5, 1000)  # a device write through the imported mapping
    ctx.sync()
    got = src.download(np.uint32, MiB)
    exp = pattern.copy()
    exp[MiB // 4: MiB // 4 + 1000] = 0xA5A5A5A5
    np.testing.assert_array_equal(got, exp)
    ctx.extmem_release(mem)
    with pytest.raises(A.BzError) as e:
        ctx.extmem_release(mem)
    assert e.value.status == A.E_INVALID
    src.free()


def test_frame_through_imported_memory(ctx):
    """Depth, commands and counts live in one exported allocation (as a Vulkan allocation with suballocated
    buffers would); the pyramid is built from the imported depth and both passes write through the import."""
    n = 262144 + 17
    sc = H.c1_scene(n)
    scene = H.as_tuple(sc)
    h = ctx.scene_create(*scene)
    depth = H.c1_depth()
    block = ctx.alloc(32 * MiB)
    off_depth, off_early, off_late, off_counts = 0, 9 * MiB, 16 * MiB, 31 * MiB
    assert depth.nbytes <= off_early and off_early + n * 24 <= off_late and off_late + n * 24 <= off_counts
    fd = export_dmabuf(block)
    mem = ctx.extmem_import(fd, block.nbytes)
    os.close(fd)
    d_view = ctx.extmem_map(mem, off_depth, depth.nbytes)
    e_view = ctx.extmem_map(mem, off_early, n * 24)
    l_view = ctx.extmem_map(mem, off_late, n * 24)
    c_view = ctx.extmem_map(mem, off_counts, 8)
    d_view.upload(depth)
    c_view.fill_u32(0xDEADBEEF, 2)
    p = ctx.pyramid_create(1024, 512, 11)
    ctx.pyramid_build(p, d_view.ptr, 1920, 1080)
    v = H.view(pyr=(1024, 512))
    vis0 = bz.synth_visibility(H.SEED_C1 + 5, n, 0.5)
    ctx.set_visibility(h, vis0)
    e_cnt = bz.MappedBuffer(ctx, c_view.ptr, 4)
    l_cnt = bz.MappedBuffer(ctx, c_view.ptr + 4, 4)
    ctx.cull(h, v, A.PASS_EARLY, e_view, e_cnt, capacity=n, flags=A.CULL_FRAME_CACHE)
    ctx.cull(h, v, A.PASS_LATE, l_view, l_cnt, capacity=n, pyramid=p, flags=A.CULL_FRAME_CACHE)
    ctx.check_kernels()
    # read back through the exporter's allocation
    counts = block.download(np.uint32, 2, offset_bytes=off_counts)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    vb = H.view_bytes(v)
    e_exp = orc.cull(scene, vb, A.PASS_EARLY, vis_in=vis0)
    vexp = vis0.copy()
    l_exp = orc.cull(scene, vb, A.PASS_LATE, opyr, vis_in=vis0, vis_out=vexp)
    assert (int(counts[0]), int(counts[1])) == (e_exp[1], l_exp[1])
    e_got = block.download(np.uint32, e_exp[1] * 6, offset_bytes=off_early).reshape(-1, 6)
    l_got = block.download(np.uint32, l_exp[1] * 6, offset_bytes=off_late).reshape(-1, 6)
    np.testing.assert_array_equal(e_got, H.cmds_u32(e_exp[0]))
    np.testing.assert_array_equal(l_got, H.cmds_u32(l_exp[0]))
    np.testing.assert_array_equal(ctx.get_visibility(h, n), vexp)
    ctx.extmem_release(mem)
    ctx.pyramid_destroy(p)
    ctx.scene_destroy(h)
    block.free()


def test_interop_argument_errors(ctx):
    h = C.c_uint64()
    lib = A.load()
    assert lib.bz_extmem_import(ctx.handle, -1, 4096, C.byref(h)) == A.E_INVALID
    assert lib.bz_extsem_import(ctx.handle, -1, A.EXTSEM_BINARY, C.byref(h)) == A.E_INVALID
    r, w = os.pipe()
    try:
        assert lib.bz_extsem_import(ctx.handle, r, 7, C.byref(h)) == A.E_INVALID  # unknown kind: no import tried
        assert lib.bz_extmem_import(ctx.handle, r, 0, C.byref(h)) == A.E_INVALID  # zero size: no import tried
    finally:
        os.close(r)
        os.close(w)
    p = C.c_void_p()
    assert lib.bz_extmem_map(ctx.handle, 12345, 0, 16, C.byref(p)) == A.E_INVALID
    assert lib.bz_extmem_release(ctx.handle, 12345) == A.E_INVALID
    assert lib.bz_extsem_wait(ctx.handle, 12345, 0, None) == A.E_INVALID
    assert lib.bz_extsem_signal(ctx.handle, 12345, 0, None) == A.E_INVALID
    assert lib.bz_extsem_release(ctx.handle, 12345) == A.E_INVALID


def test_map_range_checked(ctx):
    src = ctx.alloc(2 * MiB)
    fd = export_dmabuf(src)
    mem = ctx.extmem_import(fd, src.nbytes)
    os.close(fd)
    p = C.c_void_p()
    lib = A.load()
    assert lib.bz_extmem_map(ctx.handle, mem, MiB, MiB + 4, C.byref(p)) == A.E_INVALID
    assert lib.bz_extmem_map(ctx.handle, mem, 2 * MiB + 1, 1, C.byref(p)) == A.E_INVALID
    assert lib.bz_extmem_map(ctx.handle, mem, 0, 0, C.byref(p)) == A.E_INVALID
    ctx.extmem_release(mem)
    src.free()
