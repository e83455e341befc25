"""The host side of bench.py's output contract (no GPU): the one-GPU guard and the committed traffic summary that
the roofline's `traffic` field is read from."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_traffic_summary_is_committed_and_consistent():
    t = bench.load_traffic()
    assert t is not None, "profiles/traffic_frame_c1.json is missing or unreadable"
    fetch = t["fetch_kib_raw"] * 1024 * t["fetch_scale"]
    write = t["write_kib_raw"] * 1024 * t["write_scale"]
    assert abs(fetch + write - t["bytes_per_launch"]) <= 0.001 * t["bytes_per_launch"]
    # the command that made it runs committed files only
    assert "bench.py" in t["command"] and "_work" not in t["command"]
    # it names the kernel sources it measured (bench.py reports it as stale when they differ from the library's)
    src = t.get("kernel_sources", "")
    assert len(src) == 16 and all(c in "0123456789abcdef" for c in src)


def test_profiles_name_their_kernel_sources_and_commands():
    for p in sorted((ROOT / "profiles").glob("*.md")):
        text = p.read_text()
        assert "Kernel sources: `" in text, p.name
        assert "rocprofv3 " in text and "_work" not in text, p.name


def test_percentile_picks_sample_values():
    xs = [5.0, 1.0, 3.0, 2.0, 4.0]
    assert bench.pct(xs, 0) == 1.0 and bench.pct(xs, 50) == 3.0 and bench.pct(xs, 100) == 5.0
    assert bench.pct([7.0], 99) == 7.0
