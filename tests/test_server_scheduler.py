"""The serving loop's scheduler (blitzen_amd/csrc/server_core.h) on the host, over a fake device: several client
threads against the launcher and completer threads, with failed launches, failed polls and given-up waits
(tests/native/server_check.cpp). Built and run twice: under ThreadSanitizer, and under AddressSanitizer + UBSan."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "native" / "server_check.cpp"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("sanitizer", ["thread", "address,undefined"])
def test_scheduler_with_fake_device(tmp_path, sanitizer):
    exe = tmp_path / "server_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-pthread", f"-fsanitize={sanitizer}",
                    "-fno-omit-frame-pointer", str(SRC), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600,
                       env={"TSAN_OPTIONS": "halt_on_error=1 second_deadlock_stack=1",
                            "ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "halt_on_error=1"})
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout[-3000:] + r.stderr[-3000:]
