"""Shared workload builders for the tests (seeded, identical for the oracle and the GPU path)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from blitzen_amd import cull as bz

SEED_C0 = 0xB1172E00
SEED_C1 = 0xB1172E01
SEED_C2 = 0xB1172E02
SEED_C4 = 0xB1172E04


def as_tuple(sc: bz.SynthScene):
    return (sc.objects, sc.transforms, sc.surfaces, sc.lods)


def c0_scene():
    """16,384 objects, 64 surfaces x 4 LODs, positions U[-500,500]^3, scale U[0.5,2]."""
    return bz.synth_scene(SEED_C0, 16384, 64, 4)


def c1_scene(n: int = 4194304, seed: int = SEED_C1):
    """Config-1 distributions: n objects over 256 surfaces x 8 LODs."""
    return bz.synth_scene(seed, n, 256, 8)


def view(fov=70.0, w=1920, h=1080, znear=0.1, zfar=650.0, pos=(0.0, 0.0, 0.0), yaw=0.0, pitch=0.0, pyr=(0, 0)):
    v = bz.camera_setup(fov, w, h, znear, zfar, pos, yaw, pitch)
    v.pyramidWidth = float(pyr[0])
    v.pyramidHeight = float(pyr[1])
    return v


def view_bytes(v) -> bytes:
    return bytes(memoryview(v))


def c1_depth(seed: int = SEED_C1, w=1920, h=1080, occluders=512, p11=None, znear=0.1):
    if p11 is None:
        p11 = view(w=w, h=h).proj5
    return bz.synth_depth(seed, w, h, occluders, p11, znear)


def cmds_u32(structured: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(structured).view(np.uint32).reshape(-1, 6)
