"""The integration guide's frame, as a C11 translation unit that includes include/bzcull.h and links libbzcull.so
(tests/native/integration_frame.c): it must compile warning-free as strict C11 against the header, and run. Without
a GPU it reports that no device is present; the GPU test compares its draw counts with the oracle's."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from blitzen_amd import _capi

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "native" / "integration_frame.c"


def build(tmp_path) -> Path:
    exe = tmp_path / "integration_frame"
    lib_dir = _capi.LIB_PATH.parent
    subprocess.run(["gcc", "-std=c11", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1", "-I", str(ROOT / "include"),
                    str(SRC), "-L", str(lib_dir), "-lbzcull", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_c11_caller_compiles_links_and_runs(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([str(exe), "1000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith(("no device", "early ")), r.stdout


@pytest.mark.gpu
def test_c11_caller_frame_matches_oracle(tmp_path):
    from blitzen_amd import cull as bz
    from oracle import pyoracle as orc

    n = 100000
    exe = build(tmp_path)
    r = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    words = r.stdout.split()
    assert words[0] == "early" and words[2] == "late", r.stdout
    early, late = int(words[1]), int(words[3])
    seed = 0xB1172E01
    sc = bz.synth_scene(seed, n, 256, 8)
    scene = (sc.objects, sc.transforms, sc.surfaces, sc.lods)
    v = bz.camera_setup(70.0, 1920, 1080, 0.1, 650.0)
    v.pyramidWidth, v.pyramidHeight = 1024.0, 512.0
    vb = bz.view_bytes(v)
    vis0 = bz.synth_visibility(seed, n, 0.5)
    depth = bz.synth_depth(seed, 1920, 1080, 512, v.proj5, 0.1)
    opyr = orc.pyramid(orc.pyramid_build(depth, 1024, 512, 11), 1024, 512, 11)
    _, e = orc.cull(scene, vb, 0, vis_in=vis0)
    vis = vis0.copy()
    _, l_ = orc.cull(scene, vb, 1, pyr=opyr, vis_in=vis, vis_out=vis)
    assert (early, late) == (e, l_) and early > 0 and late > 0
    assert np.count_nonzero(vis) > 0
