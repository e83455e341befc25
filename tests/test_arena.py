"""The serving loop's HBM-budget allocator (host logic, no GPU): built with g++ and checked against a brute-force
first-fit model (tests/native/arena_check.cpp)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_arena_first_fit_model(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined", str(HERE / "native" /
                    "arena_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() == "ok", out.stdout
