"""Benchmark of the GPU-driven draw-cull path on one MI355X.

Workload (one step = one frame of the two-pass cull): 4,194,304 objects over 256 surfaces x 8 LODs, positions
U[-500,500]^3, scale U[0.5,2], unit quaternions; camera at the origin looking +z, 70 deg fov, 1920x1080, reverse-z,
zNear 0.1, zFar 650; a synthetic 1920x1080 reverse-z depth buffer (ground plane y = -10 plus 512 seeded occluder
rectangles of depth U[0.001,0.2]) reduced to a 1024x512-based 11-mip min-depth pyramid; last-frame visibility
Bernoulli(0.5). A step runs: pyramid build -> early pass (frustum + LOD over last-frame-visible objects) -> late
pass (frustum + Hi-Z over all objects, emits newly visible, writes visibility). Both passes use the frame cache:
the early launch runs the frustum test (and LOD) for every object and records the survivors, and the late launch
reads only those; the commands, counts and visibility are identical to the full passes.

Inputs are resident in HBM before timing. Steps rotate over --copies independent copies of the scene, depth buffer,
pyramid and visibility buffers (4 x ~200 MB by default) so no step finds its inputs in the 256 MB MALL left by the
previous one. The late pass writes visibility out of place so every step does identical work.

Prints ONE JSON line (rank 0). object-culls/s = objects / (time of one full frame).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N_OBJECTS = 4194304
N_SURFACES, N_LODS = 256, 8
DEPTH_W, DEPTH_H, OCCLUDERS = 1920, 1080, 512
PYR_W, PYR_H, PYR_MIPS = 1024, 512, 11
SEED = 0xB1172E01
HBM_PEAK_GBS = 8000.0
LATE_BYTES_PER_OBJECT = 48  # RenderObject 8 + MeshTransform 32 + visibility read 4 + write 4
EARLY_BYTES_PER_OBJECT = 4  # visibility read; + 40 B per last-frame-visible object
CMD_BYTES = 24


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def ensure_built():
    from blitzen_amd import _capi

    if not _capi.LIB_PATH.exists():
        import __graft_entry__

        __graft_entry__.build()


def cpu_baseline(budget_s: float):
    """The oracle (OpenMP port of the reference shaders) on the host's cores, full C1 frame: one thread per physical
    core available to this process (OMP_PROC_BIND=close, OMP_PLACES=cores) and one thread, each the median of >= 10
    runs after 2 warm-ups, in a child process (oracle/cpu_frame.py) so that the binding applies to the OpenMP runtime
    there and not to this GPU process."""
    import subprocess

    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_frame", "--seconds", str(budget_s)], cwd=str(ROOT),
                       env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline failed ({r.returncode}): {r.stderr[-2000:]}")
    d = json.loads(r.stdout.strip().splitlines()[-1])
    m, one = d["all_cores"], d["single_thread"]
    return {"value": m["object_culls_per_s"], "unit": "object-culls/s", "cores": m["threads"], "kind": "port",
            "sample": f"full C1 frame (pyramid + early + late, {N_OBJECTS} objects), median of {m['runs']} runs after "
                      f"2 warm-ups, OpenMP {m['threads']} threads on {m['threads']} physical cores "
                      f"(OMP_PROC_BIND=close OMP_PLACES=cores), {d['cpu_model']}",
            "ms_per_frame": m["ms_per_frame"],
            "single_thread": {"value": one["object_culls_per_s"], "ms_per_frame": one["ms_per_frame"],
                              "runs": one["runs"]},
            "cpu_model": d["cpu_model"], "nproc": d["nproc"], "physical_cores_available": d["physical_cores_available"],
            "cgroup_cpu_quota": d["cgroup_cpu_quota"], "omp_num_threads_env": d["omp_num_threads_env"]}


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))]


def bench_batched(ctx, stream, steps: int, warmup: int):
    """256 views of uneven size: one captured graph (pyramid builds + batched cull) vs per-view launches."""
    import numpy as np

    from blitzen_amd import workloads as W
    from blitzen_amd.hip import Event

    t0 = time.perf_counter()
    plan = W.batched_plan()
    bv = W.BatchedViews.build(ctx, plan, stream=stream)
    log(f"bench: batched workload ({len(plan)} views, {bv.n_objects} object-view pairs) built in "
        f"{time.perf_counter() - t0:.1f}s")
    # the autotuner keeps its choice in the context: the single-view tuning it picks for these views (used by the
    # per-view baseline) is put back afterwards, so the legs that follow run with the built-in one as the frame did
    single_view = ctx.get_tuning(0)
    tuned = ctx.autotune(1, bv.jobs, stream=stream)
    tuned0 = ctx.autotune(0, bv.jobs, stream=stream)
    graph = ctx.batch_graph_create(bv.pyr_jobs, bv.jobs)

    def timed(fn):
        for _ in range(warmup):
            fn()
        ctx.sync(stream)
        evs = [(Event(), Event()) for _ in range(steps)]
        t_start = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            fn()
            b.record(stream)
        ctx.sync(stream)
        wall = (time.perf_counter() - t_start) / steps
        ctx.check_kernels(stream)
        return [a.elapsed_ms(b) for a, b in evs], wall

    t_graph, wall_graph = timed(lambda: ctx.graph_launch(graph, stream))
    c_graph = bv.counts(stream).copy()
    t_plain, wall_plain = timed(lambda: bv.run_per_view(stream))
    c_plain = bv.counts(stream)
    if not np.array_equal(c_graph, c_plain):
        raise RuntimeError("batched and per-view draw counts differ")
    ctx.graph_destroy(graph)
    ctx.set_tuning(0, single_view.block, single_view.items, single_view.grid_per_cu)
    n_views = len(plan)
    cmds = int(c_graph.sum())
    alg = W.algorithmic_bytes_batched(plan, cmds)
    g_ms, p_ms = statistics.median(t_graph), statistics.median(t_plain)
    out = {
        "workload": f"{n_views} views, per-view objects log-uniform in [1024, 1048576] ({bv.n_objects} object-view "
                    f"pairs), per-view depth 512x512 / 1024x1024 / 2048x1024 with 64-1024 occluders, per-view "
                    f"pyramid built in the same launch sequence, late-pass semantics",
        "views": n_views,
        "object_view_pairs": bv.n_objects,
        "views_per_s": round(n_views / (g_ms * 1e-3), 1),
        "object_culls_per_s": round(bv.n_objects / (g_ms * 1e-3), 1),
        "graph_ms_p50": round(g_ms, 4),
        "graph_ms_p99": round(pct(t_graph, 99), 4),
        "graph_wall_ms": round(wall_graph * 1e3, 4),
        "achieved_gbs": round(alg / (g_ms * 1e-3) / 1e9, 1),
        "algorithmic_bytes": alg,
        "per_view_launches": {"views_per_s": round(n_views / (p_ms * 1e-3), 1), "ms_p50": round(p_ms, 4),
                              "ms_p99": round(pct(t_plain, 99), 4), "wall_ms": round(wall_plain * 1e3, 4)},
        "speedup_vs_per_view": round(p_ms / g_ms, 2),
        "tuning": {"batched": [tuned.block, tuned.items], "single_view": [tuned0.block, tuned0.items,
                                                                          tuned0.grid_per_cu]},
        "commands": cmds,
    }
    bv.free()
    return out


def bench_served(ctx, rates, duration_s: float):
    """Poisson request streams against the serving loop over 8 resident scenes. Latency is measured from each
    request's scheduled arrival to its completion, so a submitter that falls behind the schedule shows up in it
    (no coordinated omission); the submitter's own delay (submission - arrival) and the latency from submission are
    reported beside it."""
    import numpy as np

    from blitzen_amd import _capi as A
    from blitzen_amd import workloads as W

    t0 = time.perf_counter()
    scenes = W.served_scenes(ctx)
    log(f"bench: served scenes ({sum(n for _, n in scenes)} objects) resident in {time.perf_counter() - t0:.1f}s")
    budget = 64 << 30
    ctx.server_start(budget, inflight=4, max_batch=64, flags=A.SERVER_RELEASE_ON_COMPLETE)
    out = {"workload": f"8 resident scenes {W.SERVE_SCENES} (positions U[-500,500]^3, scales U[0.5,2], 256 surfaces "
                       f"x 8 LODs); Poisson arrivals; each request: seeded camera, depth "
                       f"{W.SERVE_DEPTH_SIZES[0]}..{W.SERVE_DEPTH_SIZES[-1]} rasterised on the device from its seed "
                       f"({W.SERVE_OCCLUDERS} occluders), own pyramid, Hi-Z cull of the whole scene; 64 GB HBM "
                       f"budget; 4 batches in flight, <= 64 requests per batch",
           "latency": "scheduled arrival -> completion (submit_delay: arrival -> submission)", "rates": []}
    try:
        warm, offs, _ = W.served_requests(scenes, 2000, 0.02, seed=1)
        tk, _ = ctx.submit_schedule(warm, offs)
        for t in tk:
            ctx.wait(int(t), 60)
            ctx.release(int(t))
        for rate in rates:
            reqs, offs, objs = W.served_requests(scenes, rate, duration_s)
            tickets, st = ctx.submit_schedule(reqs, offs)
            lat, lat_sub, delay, arr, comp, done_objs, bsz, bad = [], [], [], [], [], 0, [], 0
            for t, n_obj in zip(tickets, objs):
                if not t:
                    bad += 1
                    continue
                r = ctx.wait(int(t), 120)
                if r.status == A.OK:
                    lat.append((r.complete_ns - r.arrival_ns) / 1e6)
                    lat_sub.append((r.complete_ns - r.submit_ns) / 1e6)
                    delay.append((r.submit_ns - r.arrival_ns) / 1e6)
                    arr.append(r.arrival_ns)
                    comp.append(r.complete_ns)
                    done_objs += int(n_obj)
                    bsz.append(r.batch_size)
                else:
                    bad += 1
                ctx.release(int(t))
            span = (max(comp) - min(arr)) / 1e9 if comp else float("nan")
            out["rates"].append({
                "offered_req_per_s": rate, "requests": len(reqs), "failed": bad,
                "achieved_req_per_s": round(len(lat) / span, 1),
                "object_culls_per_s": round(done_objs / span, 1),
                "latency_ms_p50": round(pct(lat, 50), 4), "latency_ms_p99": round(pct(lat, 99), 4),
                "latency_ms_max": round(max(lat), 4),
                "submit_delay_ms_p50": round(pct(delay, 50), 4), "submit_delay_ms_p99": round(pct(delay, 99), 4),
                "latency_from_submit_ms_p50": round(pct(lat_sub, 50), 4),
                "latency_from_submit_ms_p99": round(pct(lat_sub, 99), 4),
                "mean_batch": round(statistics.mean(bsz), 2)})
            log(f"bench: served {rate}/s -> p50 {out['rates'][-1]['latency_ms_p50']} ms, "
                f"p99 {out['rates'][-1]['latency_ms_p99']} ms, {out['rates'][-1]['achieved_req_per_s']} req/s")
        st = ctx.server_stats()
        out["budget_peak_bytes"] = int(st.budget_peak)
    finally:
        ctx.server_stop()
        for h, _ in scenes:
            ctx.scene_destroy(h)
    return out


def bench_clusters(ctx, stream, steps: int, warmup: int):
    """Cluster mode: per frame, 1% transform updates then the two-level cluster cull."""
    from blitzen_amd import workloads as W
    from blitzen_amd.hip import Event

    t0 = time.perf_counter()
    cf = W.ClusterFrames.build(ctx, stream=stream)
    ups = [cf.updates(f) for f in range(warmup + steps)]
    log(f"bench: cluster scene built in {time.perf_counter() - t0:.1f}s")
    for f in range(warmup):
        cf.frame(*ups[f], stream)
    ctx.sync(stream)
    evs = [(Event(), Event()) for _ in range(steps)]
    t_start = time.perf_counter()
    for k, (a, b) in enumerate(evs):
        a.record(stream)
        cf.frame(*ups[warmup + k], stream)
        b.record(stream)
    ctx.sync(stream)
    wall = (time.perf_counter() - t_start) / steps
    ctx.check_kernels(stream)
    records, draws = cf.result_counts(stream)
    ms = statistics.median([a.elapsed_ms(b) for a, b in evs])
    # the cluster cull alone (no transform upload): one event pair around `steps` calls
    c0, c1 = Event(), Event()
    c0.record(stream)
    for _ in range(steps):
        cf.cull(stream)
    c1.record(stream)
    ctx.sync(stream)
    ctx.check_kernels(stream)
    cull_ms = c0.elapsed_ms(c1) / steps
    # algorithmic bytes of the cull: object pass 40 B per object (RenderObject 8 + transform 32), per record 12 B
    # written and 12 B read back + 20 B of cluster data (sphere, cone), 24 B per draw, two counts
    cull_bytes = cf.n_objects * 40 + records * (12 + 12 + 20) + draws * 24 + 8
    out = {"workload": "1,048,576 objects over 128 surfaces x 8 LODs with 16-128 clusters of 64 triangles (LOD k "
                       "owns max(1, c >> 2k)); per frame 1% seeded transform updates (10,485 x 32 B uploaded) + object "
                       "pass (frustum + LOD, as PreClusterDrawCull) + expansion + cluster frustum/Hi-Z/cone tests "
                       "against a 1024x512 11-mip pyramid",
           "frame_ms_p50": round(ms, 4), "frame_wall_ms": round(wall * 1e3, 4),
           "objects_per_s": round(cf.n_objects / (ms * 1e-3), 1),
           "cluster_records": records, "cluster_draws": draws,
           "clusters_per_s": round(records / (ms * 1e-3), 1),
           "cull_us": round(cull_ms * 1e3, 2), "cull_algorithmic_bytes": cull_bytes,
           "cull_achieved_gbs": round(cull_bytes / (cull_ms * 1e-3) / 1e9, 1),
           "cull_records_per_s": round(records / (cull_ms * 1e-3), 1)}
    cf.free()
    return out


def bench_next_rows(ctx, stream, scenes, view, transforms, steps: int):
    """The scope table's next rows at the C1 size: the transparent (PLAIN) and ONPC passes, the LOD-instanced cull
    and 1% transform streaming. Each is timed by a HIP event pair around `steps` launches on the launch stream,
    rotating over the resident scene copies (so the 160 MB scene is not L2/MALL-resident between launches)."""
    import numpy as np

    from blitzen_amd import _capi as A
    from blitzen_amd import cull as bz
    from blitzen_amd.hip import Event

    def timed(fn):
        for k in range(2 * len(scenes) + 1):  # every scene copy's lazily allocated scratch and staging slots exist
            fn(k)
        ctx.sync(stream)
        e0, e1 = Event(), Event()
        t0 = time.perf_counter()
        e0.record(stream)
        for k in range(steps):
            fn(k)
        e1.record(stream)
        ctx.sync(stream)
        wall = (time.perf_counter() - t0) / steps
        ctx.check_kernels(stream)
        return e0.elapsed_ms(e1) / steps, wall

    def row(ms, alg, extra):
        gbs = alg / (ms * 1e-3) / 1e9
        return {"us": round(ms * 1e3, 2), "object_culls_per_s": round(N_OBJECTS / (ms * 1e-3), 1),
                "algorithmic_bytes": alg, "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), **extra}

    out = {}
    cmds = ctx.alloc(N_OBJECTS * CMD_BYTES)
    cnt = ctx.alloc(4)
    for name, pass_ in (("plain", A.PASS_PLAIN), ("onpc", A.PASS_ONPC)):
        ms, _ = timed(lambda k: ctx.cull(scenes[k % len(scenes)], view, pass_, cmds, cnt, capacity=N_OBJECTS,
                                         stream=stream))
        n = int(cnt.download(np.uint32, 1, stream)[0])
        out[name] = row(ms, N_OBJECTS * 40 + n * CMD_BYTES + 4, {"commands": n})
    # LOD-instanced: every LOD's list has room for 8192 instances (the largest list here holds a few thousand)
    n_lods = N_SURFACES * N_LODS
    per_lod = 8192
    ids, counters = ctx.alloc(n_lods * per_lod * 4), ctx.alloc(n_lods * 8)
    icmds, icnt = ctx.alloc(n_lods * 32), ctx.alloc(4)
    ms, _ = timed(lambda k: ctx.cull_instanced(scenes[k % len(scenes)], view, per_lod, ids, counters, icmds, icnt,
                                               stream=stream))
    k_cmds = int(icnt.download(np.uint32, 1, stream)[0])
    ctr = counters.download(np.uint32, n_lods * 2, stream).reshape(-1, 2)
    inst = int(ctr[:, 1].sum())
    out["instanced"] = row(ms, N_OBJECTS * 40 + inst * 4 + n_lods * 8 + k_cmds * 32 + 4,
                           {"instances": inst, "lod_commands": k_cmds, "instances_per_lod": per_lod,
                            "largest_list": int(ctr[:, 1].max())})
    # transform streaming: 1% of the transforms per call (ids + 32-byte transforms from host arrays)
    n_upd = N_OBJECTS // 100
    upd = [bz.synth_transform_updates(SEED, f, transforms, n_upd) for f in range(4)]
    ms, wall = timed(lambda k: ctx.update_transforms(scenes[k % len(scenes)], upd[k % 4][0], upd[k % 4][1], stream))
    host_bytes = n_upd * (4 + 32)
    out["transform_updates"] = {"updates_per_call": n_upd, "us": round(ms * 1e3, 2), "wall_us": round(wall * 1e6, 2),
                                "host_bytes": host_bytes, "pcie_inclusive_gbs": round(host_bytes / wall / 1e9, 2),
                                "updates_per_s": round(n_upd / wall, 1)}
    # the same updates written by the caller straight into the pinned staging slots (filled once per slot before the
    # timed calls; each call hands out the next slot and commits it: two uploads and the device scatter)
    h0 = scenes[0]
    for k in range(2):
        sid, sxf = ctx.transform_staging(h0, n_upd)
        sid[:] = upd[k][0]
        sxf[:] = upd[k][1]
        ctx.commit_transforms(h0, n_upd, stream)

    def staged(k):
        ctx.transform_staging(h0, n_upd)
        ctx.commit_transforms(h0, n_upd, stream)

    ms, wall = timed(staged)
    out["transform_updates_staged"] = {"updates_per_call": n_upd, "us": round(ms * 1e3, 2),
                                       "wall_us": round(wall * 1e6, 2), "host_bytes": host_bytes,
                                       "pcie_inclusive_gbs": round(host_bytes / wall / 1e9, 2),
                                       "updates_per_s": round(n_upd / wall, 1)}
    for b in (cmds, cnt, ids, counters, icmds, icnt):
        b.free()
    return out


def bench_size_scaling(ctx, stream, view, steps: int, copies: int = 2):
    """The PLAIN pass of one job at 4, 8 and 16 M objects (ranges [0, n) of two resident 16 M-object scenes, used
    in turn so that a range is not cache-resident between launches). The pass is decide + emit; a job of more than
    4096 tiles also scans its tile counts, so that every emit workgroup reads its output offset instead of re-summing
    the earlier tiles' counts. Time per million objects stays flat when the emit is linear in the object count."""
    import numpy as np

    from blitzen_amd import _capi as A
    from blitzen_amd import cull as bz
    from blitzen_amd.hip import Event

    n_max = 16 * 1048576
    sc = bz.synth_scene(SEED + 16, n_max, N_SURFACES, N_LODS)
    scenes = [ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods) for _ in range(copies)]
    del sc
    cmds, cnt = ctx.alloc(n_max // 2 * CMD_BYTES), ctx.alloc(4)
    out = {"workload": "PLAIN pass, one job of n objects (C1 distributions), rotating over "
                       f"{copies} resident 16 M-object scenes", "sizes": []}
    for n in (4 * 1048576, 8 * 1048576, n_max):
        def fn(k):
            ctx.cull(scenes[k % copies], view, A.PASS_PLAIN, cmds, cnt, capacity=n_max // 2, n=n, stream=stream)
        for k in range(2 * copies):
            fn(k)
        ctx.sync(stream)
        e0, e1 = Event(), Event()
        e0.record(stream)
        for k in range(steps):
            fn(k)
        e1.record(stream)
        ctx.sync(stream)
        ctx.check_kernels(stream)
        us = e0.elapsed_ms(e1) / steps * 1e3
        n_cmd = int(cnt.download(np.uint32, 1, stream)[0])
        alg = n * 40 + n_cmd * CMD_BYTES + 4
        out["sizes"].append({"objects": n, "tiles": n // 1024, "scanned_prefix": n // 1024 > 4096,
                             "us": round(us, 2), "us_per_mobject": round(us / (n / 1048576), 3), "commands": n_cmd,
                             "achieved_gbs": round(alg / (us * 1e-6) / 1e9, 1)})
    for h in scenes:
        ctx.scene_destroy(h)
    cmds.free()
    cnt.free()
    return out


def load_traffic():
    p = ROOT / "profiles" / "traffic_frame_c1.json"
    if p.exists():
        try:
            return json.loads(p.read_text())
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--copies", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-clusters", action="store_true")
    ap.add_argument("--no-scaling", action="store_true", help="skip the PLAIN pass at 4, 8 and 16 M objects")
    ap.add_argument("--no-graph-replay", action="store_true", help="skip timing the frames replayed from hipGraphs")
    ap.add_argument("--no-next", action="store_true", help="skip the PLAIN / ONPC / instanced / transform-update leg")
    ap.add_argument("--served-seconds", type=float, default=0.5)
    ap.add_argument("--no-frame-cache", action="store_true",
                    help="run the C1 passes without BZ_CULL_FRAME_CACHE (the LATE pass reads every object)")
    ap.add_argument("--calib-mib", type=int, default=0,
                    help="first copy N MiB device to device 3 times (counter calibration for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    ensure_built()
    import numpy as np

    from blitzen_amd import _capi as A
    from blitzen_amd import cull as bz
    from blitzen_amd.hip import Event

    t0 = time.perf_counter()
    sc = bz.synth_scene(SEED, N_OBJECTS, N_SURFACES, N_LODS)
    v = bz.camera_setup(70.0, DEPTH_W, DEPTH_H, 0.1, 650.0)
    v.pyramidWidth, v.pyramidHeight = float(PYR_W), float(PYR_H)
    depth = bz.synth_depth(SEED, DEPTH_W, DEPTH_H, OCCLUDERS, v.proj5, 0.1)
    vis0 = bz.synth_visibility(SEED, N_OBJECTS, 0.5)
    log(f"bench: workload generated in {time.perf_counter() - t0:.1f}s")

    ctx = bz.Context(0)
    name, ncu, hbm = ctx.device_info()
    log(f"bench: device {name}, {ncu} CUs, {hbm / 2**30:.0f} GiB")
    stream = ctx.stream()
    copies = []
    for c in range(args.copies):
        h = ctx.scene_create(sc.objects, sc.transforms, sc.surfaces, sc.lods)
        vin = ctx.alloc(N_OBJECTS * 4)
        vin.upload(vis0)
        vout = ctx.alloc(N_OBJECTS * 4)
        dbuf = ctx.alloc(depth.nbytes)
        dbuf.upload(depth)
        pyr = ctx.pyramid_create(PYR_W, PYR_H, PYR_MIPS)
        copies.append((h, vin, vout, dbuf, pyr))
    cmds_e = ctx.alloc(N_OBJECTS * CMD_BYTES)
    cmds_l = ctx.alloc(N_OBJECTS * CMD_BYTES)
    cnt = ctx.alloc(16)
    cnt_e_ptr, cnt_l_ptr = cnt.ptr, cnt.ptr + 4

    class _Cnt:  # count sub-buffers
        def __init__(self, p):
            self.ptr = p

    ce, cl = _Cnt(cnt_e_ptr), _Cnt(cnt_l_ptr)
    fc_flag = 0 if args.no_frame_cache else A.CULL_FRAME_CACHE

    # Each copy's two bz_cull argument structs are built once and the three calls per frame go straight to the
    # library, as a native renderer would make them: the per-call Python work of the convenience wrappers (building
    # ctypes structs) took enough host time that on some hosts the GPU waited between launches.
    import ctypes as C

    L = bz.lib()
    vptr = C.pointer(v)
    frame_args = []
    for h, vin, vout, dbuf, pyr in copies:
        ae = A.CullArgs(scene=h, pyramid=0, pass_=A.PASS_EARLY, flags=fc_flag, view=vptr, cmds_dev=cmds_e.ptr,
                        count_dev=cnt_e_ptr, capacity=N_OBJECTS, object_begin=0, object_count=0,
                        visibility_in_dev=vin.ptr, visibility_out_dev=None)
        al = A.CullArgs(scene=h, pyramid=pyr, pass_=A.PASS_LATE, flags=fc_flag, view=vptr, cmds_dev=cmds_l.ptr,
                        count_dev=cnt_l_ptr, capacity=N_OBJECTS, object_begin=0, object_count=0,
                        visibility_in_dev=vin.ptr, visibility_out_dev=vout.ptr)
        frame_args.append((ae, al, C.byref(ae), C.byref(al)))

    def step(k, ev=None):
        h, vin, vout, dbuf, pyr = copies[k % len(copies)]
        _, _, ae, al = frame_args[k % len(copies)]
        if ev:
            ev[0].record(stream)
        s = L.bz_pyramid_build(ctx.handle, pyr, dbuf.ptr, DEPTH_W, DEPTH_H, stream)
        if s:
            ctx.check(s, "bz_pyramid_build")
        if ev:
            ev[1].record(stream)
        s = L.bz_cull(ctx.handle, ae, stream)
        if s:
            ctx.check(s, "bz_cull (EARLY)")
        if ev:
            ev[2].record(stream)
        s = L.bz_cull(ctx.handle, al, stream)
        if s:
            ctx.check(s, "bz_cull (LATE)")
        if ev:
            ev[3].record(stream)

    if args.calib_mib:
        a, b = ctx.alloc(args.calib_mib << 20), ctx.alloc(args.calib_mib << 20)
        for _ in range(3):
            ctx.check(bz.lib().bz_memcpy_d2d(ctx.handle, b.ptr, a.ptr, args.calib_mib << 20, stream), "copy")
        ctx.sync(stream)
        a.free()
        b.free()
    for k in range(args.warmup):
        step(k)
    ctx.check_kernels(stream)
    counts = cnt.download(np.uint32, 2, stream)
    n_early, n_late = int(counts[0]), int(counts[1])

    # timed region: wall clock over exactly K steps, device synchronized on both sides; a HIP event pair on the
    # launch stream brackets the same K frames (the roofline's average frame duration)
    ctx.sync(stream)
    e_start, e_end = Event(), Event()
    t_start = time.perf_counter()
    e_start.record(stream)
    for k in range(args.steps):
        step(k)
    e_end.record(stream)
    enqueue = time.perf_counter() - t_start  # host time to issue the K frames (close to `wall`: host-bound)
    ctx.sync(stream)
    wall = time.perf_counter() - t_start
    ctx.check_kernels(stream)
    frame_ms = e_start.elapsed_ms(e_end) / args.steps

    # per-pass durations with HIP events between the passes (separate run of the same work; the markers between
    # passes add a little time of their own)
    evs = [[Event() for _ in range(4)] for _ in range(args.steps)]
    for k in range(args.steps):
        step(k, evs[k])
    ctx.sync(stream)
    t_pyr = [e[0].elapsed_ms(e[1]) for e in evs]
    t_early = [e[1].elapsed_ms(e[2]) for e in evs]
    t_late = [e[2].elapsed_ms(e[3]) for e in evs]
    t_frame = [e[0].elapsed_ms(e[3]) for e in evs]

    # The same frames again, each copy's frame (pyramid build + EARLY + LATE) captured once in a hipGraph and
    # replayed with one launch per frame (reported beside the headline: the graph's kernel nodes start later than
    # back-to-back launches on the stream, so the replayed frame is slower here).
    graph_replay = None
    if not args.no_graph_replay:
        from blitzen_amd.hip import StreamGraph

        graphs = [StreamGraph(stream, lambda k=k: step(k)) for k in range(len(copies))]
        for k in range(args.warmup):
            graphs[k % len(graphs)].launch(stream)
        ctx.sync(stream)
        g_start, g_end = Event(), Event()
        t_g = time.perf_counter()
        g_start.record(stream)
        for k in range(args.steps):
            graphs[k % len(graphs)].launch(stream)
        g_end.record(stream)
        ctx.sync(stream)
        wall_g = time.perf_counter() - t_g
        ctx.check_kernels(stream)
        counts_g = cnt.download(np.uint32, 2, stream)
        if (int(counts_g[0]), int(counts_g[1])) != (n_early, n_late):
            raise RuntimeError(f"replayed frames gave counts {list(counts_g)}, eager frames {n_early}, {n_late}")
        del graphs
        graph_replay = {"ms_per_step": round(wall_g / args.steps * 1e3, 4),
                        "frame_timed_region_us": round(g_start.elapsed_ms(g_end) / args.steps * 1e3, 2),
                        "note": "each copy's frame captured once in a hipGraph, one launch per frame"}

    ms_per_step = wall / args.steps * 1e3
    value = N_OBJECTS / (wall / args.steps)
    late_ms = statistics.mean(t_late)
    late_bytes = N_OBJECTS * LATE_BYTES_PER_OBJECT + n_late * CMD_BYTES + 4
    early_bytes = N_OBJECTS * EARLY_BYTES_PER_OBJECT + int(vis0.sum()) * 40 + n_early * CMD_BYTES + 4
    pyr_bytes = DEPTH_W * DEPTH_H * 4 + sum(max(1, PYR_W >> L) * max(1, PYR_H >> L) for L in range(PYR_MIPS)) * 4
    frame_bytes = pyr_bytes + early_bytes + late_bytes
    # roofline unit: the whole frame. With the frame cache the EARLY launch does part of the LATE pass's work
    # (every object's frustum test), so per-pass figures no longer divide a pass's own bytes by its own time.
    achieved = frame_bytes / (frame_ms * 1e-3) / 1e9
    achieved_wall = frame_bytes / (ms_per_step * 1e-3) / 1e9
    traffic = load_traffic()
    from blitzen_amd.build import kernel_sources_digest

    sources = kernel_sources_digest()
    traffic_ok = bool(traffic) and not args.no_frame_cache and traffic.get("kernel_sources") == sources
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "achieved_wall": round(achieved_wall, 1), "frac_wall": round(achieved_wall / HBM_PEAK_GBS, 4),
                "traffic": traffic.get("bytes_per_launch") if traffic_ok else None,
                "kernel_sources": sources,
                "kernel": ("C1 frame: pyramid_kernel + decide_kernel (EARLY) + emit_kernel + decide_kernel (LATE) + "
                           "emit_kernel" if args.no_frame_cache else
                           "C1 frame: pyramid_kernel + decide_kernel (EARLY, frame cache) + emit_kernel + "
                           "late_cached_kernel + emit_kernel"),
                "algorithmic_bytes": frame_bytes, "avg_us": round(frame_ms * 1e3, 2)}
    if traffic and not traffic_ok and not args.no_frame_cache:  # measured on other kernels: reported, not used
        roofline["traffic_stale"] = {"bytes_per_launch": traffic.get("bytes_per_launch"),
                                     "measured_sources": traffic.get("kernel_sources")}

    result = {
        "metric": "object-culls/sec (4M objs, 1 view) + batched views/sec, p50/p99 latency",
        "value": round(value, 1),
        "unit": "object-culls/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": "C1: 4,194,304 objects, 256 surfaces x 8 LODs, 1920x1080 depth -> 1024x512 "
                               "11-mip pyramid, visibility Bernoulli(0.5), pyramid + early + late per step",
                   "objects": N_OBJECTS, "copies_rotated": args.copies, "frame_cache": not args.no_frame_cache},
        "passes_us": {"pyramid": round(statistics.mean(t_pyr) * 1e3, 2),
                      "early": round(statistics.mean(t_early) * 1e3, 2),
                      "late": round(late_ms * 1e3, 2),
                      "frame_events": round(statistics.mean(t_frame) * 1e3, 2),
                      "frame_timed_region": round(frame_ms * 1e3, 2)},
        "pass_gbs": {"pyramid": round(pyr_bytes / (statistics.mean(t_pyr) * 1e-3) / 1e9, 1),
                     "early": round(early_bytes / (statistics.mean(t_early) * 1e-3) / 1e9, 1),
                     "late": round(late_bytes / (late_ms * 1e-3) / 1e9, 1)},
        "commands": {"early": n_early, "late": n_late},
        "roofline": roofline,
        "frames": "three library calls per frame (pyramid build, EARLY, LATE), argument structs built once",
        "host_enqueue_us_per_step": round(enqueue / args.steps * 1e6, 2),
        "graph_replay": graph_replay,
    }
    if not args.no_batched:
        log("bench: batched views ...")
        result["batched"] = bench_batched(ctx, stream, max(5, min(args.steps, 20)), 2)
    if not args.no_clusters:
        log("bench: cluster mode ...")
        result["clusters"] = bench_clusters(ctx, stream, max(5, min(args.steps, 20)), 2)
    if not args.no_served:
        log("bench: served requests ...")
        result["served"] = bench_served(ctx, (2000, 8000, 32000), args.served_seconds)
    if not args.no_next:  # last: it rewrites transforms of the resident scene copies
        log("bench: next rows (plain, onpc, instanced, transform updates) ...")
        result["next_rows"] = bench_next_rows(ctx, stream, [c[0] for c in copies], v, sc.transforms,
                                              max(5, min(args.steps, 20)))
    if not args.no_scaling:
        log("bench: size scaling (PLAIN pass, 4-16 M objects) ...")
        result["size_scaling"] = bench_size_scaling(ctx, stream, v, max(5, min(args.steps, 20)))
    if not args.no_cpu:
        log("bench: CPU baseline (oracle) ...")
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
