"""CPU baseline of the C1 frame: the oracle (OpenMP C restatement of the reference shaders) timed on the host cores.

TEST INFRASTRUCTURE ONLY: run by bench.py's cpu_baseline leg as a child process, so that the OpenMP binding
(OMP_PROC_BIND=close, OMP_PLACES=cores, set before the OpenMP runtime starts) applies to it and not to the GPU
process. Runs the frame (pyramid build, EARLY pass, LATE pass) on 4,194,304 objects with one thread per physical core
available to the process, and with one thread; median of >= 10 timed runs after 2 warm-up runs each (more, up to 30,
while the time budget lasts). Prints one JSON object.

    python -m oracle.cpu_frame [--seconds S] [--threads N]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
N_OBJECTS = 4194304
SEED = 0xB1172E01


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def lscpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return cpu_model()


def physical_cores(cpus: set[int]) -> int:
    """Distinct (package, core) pairs among the logical CPUs this process may run on."""
    pairs = set()
    for c in cpus:
        base = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            pairs.add((base.joinpath("physical_package_id").read_text().strip(),
                       base.joinpath("core_id").read_text().strip()))
        except OSError:
            pairs.add(("?", str(c)))
    return len(pairs)


def cgroup_cpus() -> float | None:
    """CPU quota of the process's cgroup (v2 cpu.max), in CPUs, if one is set."""
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if quota != "max":
            return int(quota) / int(period)
    except (OSError, ValueError):
        pass
    return None


def host_info() -> dict:
    cpus = os.sched_getaffinity(0)
    quota = cgroup_cpus()
    phys = physical_cores(cpus)
    cap = os.environ.get("OMP_NUM_THREADS")
    use = phys
    if quota:
        use = min(use, max(1, int(quota)))
    if cap and cap.isdigit() and int(cap) > 0:
        use = min(use, int(cap))
    return {"cpu_model": lscpu_model(), "nproc": len(cpus), "physical_cores_available": phys,
            "cgroup_cpu_quota": quota, "omp_num_threads_env": cap, "threads_all_cores": use}


def frame_workload():
    sys.path.insert(0, str(ROOT))
    from blitzen_amd import cull as bz

    sc = bz.synth_scene(SEED, N_OBJECTS, 256, 8)
    v = bz.camera_setup(70.0, 1920, 1080, 0.1, 650.0)
    v.pyramidWidth, v.pyramidHeight = 1024.0, 512.0
    depth = bz.synth_depth(SEED, 1920, 1080, 512, v.proj5, 0.1)
    vis0 = bz.synth_visibility(SEED, N_OBJECTS, 0.5)
    return (sc.objects, sc.transforms, sc.surfaces, sc.lods), bytes(memoryview(v)), depth, vis0


def time_frames(scene, vb, depth, vis0, threads: int, budget_s: float) -> dict:
    from oracle import pyoracle as orc

    def frame():
        texels = orc.pyramid_build(depth, 1024, 512, 11, threads=threads)
        pyr = orc.pyramid(texels, 1024, 512, 11)
        orc.cull(scene, vb, 0, vis_in=vis0, threads=threads)
        vout = vis0.copy()
        orc.cull(scene, vb, 1, pyr=pyr, vis_in=vis0, vis_out=vout, threads=threads)

    for _ in range(2):
        frame()
    times = []
    t_start = time.perf_counter()
    while len(times) < 10 or (time.perf_counter() - t_start < budget_s and len(times) < 30):
        t0 = time.perf_counter()
        frame()
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"threads": threads, "runs": len(times), "ms_per_frame": round(med * 1e3, 3),
            "object_culls_per_s": round(N_OBJECTS / med, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seconds", type=float, default=20.0, help="time budget of the timed runs, both thread counts")
    ap.add_argument("--threads", type=int, default=0, help="threads of the all-cores run (0: one per core)")
    args = ap.parse_args()
    info = host_info()
    n = args.threads or info["threads_all_cores"]
    scene, vb, depth, vis0 = frame_workload()
    multi = time_frames(scene, vb, depth, vis0, n, args.seconds * 0.3)
    single = time_frames(scene, vb, depth, vis0, 1, args.seconds * 0.3)
    print(json.dumps({**info, "binding": {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES")},
                      "all_cores": multi, "single_thread": single}), flush=True)


if __name__ == "__main__":
    main()
