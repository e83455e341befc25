"""Minimal HIP runtime plumbing over ctypes (events for timing on a given stream, stream capture into a graph).

libbzcull.so links libamdhip64.so.7; loading the same soname here returns that already-loaded runtime, so events and
streams are shared with the library.
"""
from __future__ import annotations

import ctypes as C

from . import _capi

_hip = None


def hip():
    global _hip
    if _hip is None:
        _capi.load()
        h = C.CDLL("libamdhip64.so.7")
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
        h.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        h.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]
        h.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
        h.hipGraphExecDestroy.argtypes = [C.c_void_p]
        h.hipGraphDestroy.argtypes = [C.c_void_p]
        _hip = h
    return _hip


def _ok(e: int, what: str):
    if e != 0:
        raise RuntimeError(f"{what} failed with hipError {e}")


class Event:
    def __init__(self):
        p = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(p)), "hipEventCreate")
        self.h = p.value

    def record(self, stream=None):
        _ok(hip().hipEventRecord(self.h, stream), "hipEventRecord")

    def sync(self):
        _ok(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(ms), self.h, later.h), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.h:
                hip().hipEventDestroy(self.h)
        except Exception:
            pass


class StreamGraph:
    """The HIP work `fn()` enqueues on `stream`, captured once (thread-local capture mode) and replayed with one
    hipGraphLaunch. The captured calls must not allocate or synchronize: run them once before capturing."""

    def __init__(self, stream, fn):
        h = hip()
        _ok(h.hipStreamBeginCapture(stream, 1), "hipStreamBeginCapture")
        g = C.c_void_p()
        try:
            fn()
        finally:
            e = h.hipStreamEndCapture(stream, C.byref(g))
        _ok(e, "hipStreamEndCapture")
        self.graph = g.value
        x = C.c_void_p()
        _ok(h.hipGraphInstantiate(C.byref(x), self.graph, None, None, 0), "hipGraphInstantiate")
        self.exec = x.value

    def launch(self, stream):
        _ok(hip().hipGraphLaunch(self.exec, stream), "hipGraphLaunch")

    def __del__(self):
        try:
            if self.exec:
                hip().hipGraphExecDestroy(self.exec)
            if self.graph:
                hip().hipGraphDestroy(self.graph)
        except Exception:
            pass


def device_sync():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
