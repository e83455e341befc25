"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m blitzen_amd.rocprof_summary trace <results.db> [--title T] [--command C]
        per-kernel dispatch count and average / min / max / total duration (kernel trace, SQLite output);
    python -m blitzen_amd.rocprof_summary traffic <fetch.csv> <write.csv> --calib-mib N [--frame] [--command C]
                                              [--json OUT]
        HBM bytes per late pass (decide_kernel in LATE mode + the emit_kernel dispatched after it), or with --frame
        per C1 frame (every dispatch from one pyramid build to the next), from a FETCH_SIZE and a WRITE_SIZE counter
        run (CSV output). Each counter is scaled by what it reads for the runtime's device-to-device copy kernel of
        N MiB in the same run (FETCH_SIZE = N MiB read, WRITE_SIZE = N MiB written).
    python -m blitzen_amd.rocprof_summary counters <a.csv> [<b.csv> ...] [--calib-mib N] [--title T] [--command C]
                                              [--out OUT]
        per kernel, the median per dispatch of every counter in the CSV counter-collection files (one counter set per
        run); FETCH_SIZE / WRITE_SIZE also in bytes, scaled by the calibration copy when --calib-mib is given.
Every summary records the digest of the kernel sources it measured (blitzen_amd.build.kernel_sources_digest).
"""
from __future__ import annotations

import argparse
import csv
import json
import re
import sqlite3
import statistics
import sys
from collections import defaultdict
from pathlib import Path

COPY_KERNEL = "__amd_rocclr_copyBuffer"
# decide_kernel<BLOCK, ITEMS, MODE, ...> with MODE 1 = LATE, as a demangled or a mangled symbol
LATE_DECIDE = re.compile(r"decide_kernel<\d+, \d+, 1,|decide_kernelILi\d+ELi\d+ELi1E")
EMIT = re.compile(r"emit_kernel\(|11emit_kernelE")


def demangle_short(name: str) -> str:
    """A readable kernel label from the mangled symbol (no external demangler)."""
    n = name.replace(".kd", "")
    if "<" in n and "(" in n:  # already demangled: keep the template, drop the parameter list
        n = n[:n.index("(")] if n.index("(") > n.index("<") else n
        return n.replace("bz::(anonymous namespace)::", "").replace("void ", "")
    m = re.search(r"\d+([a-z_]+_kernel)", n)  # mangled: <length><name> of the kernel
    keys = (m.group(1),) if m else ()
    for key in keys:
        if key in n:
            i = n.index(key) + len(key)
            args = n[i:]
            if args.startswith("I"):
                vals = []
                for tok in args[1:].split("E"):
                    if tok.startswith(("Li", "Lj", "Lb")):  # int, unsigned, bool template arguments
                        vals.append(tok[2:])
                    else:
                        break
                return f"{key}<{','.join(vals)}>"
            return key
    return n


def sources_digest() -> str:
    from blitzen_amd.build import kernel_sources_digest

    return kernel_sources_digest()


def trace(db: str, title: str, command: str) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select s.kernel_name, d.start, d.end from rocpd_kernel_dispatch d "
                     "join rocpd_info_kernel_symbol s on d.kernel_id = s.id").fetchall()
    agg = defaultdict(list)
    for name, st, en in rows:
        agg[name].append((en - st) / 1000.0)
    lines = [f"# {title}", "", f"Kernel sources: `{sources_digest()}` (blitzen_amd.build.kernel_sources_digest)", "",
             "Command:", "", "```", command, "```", "",
             "| kernel | dispatches | avg us | min us | max us | total us | symbol |", "|---|---|---|---|---|---|---|"]
    for name, ds in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"| {demangle_short(name)} | {len(ds)} | {statistics.mean(ds):.2f} | {min(ds):.2f} | "
                     f"{max(ds):.2f} | {sum(ds):.1f} | `{name[:90]}` |")
    return "\n".join(lines) + "\n"


def read_counters(path: str):
    """{dispatch_id: (kernel_name, value)} for the single counter of a CSV counter-collection file."""
    out = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            did = int(row["Dispatch_Id"])
            val = float(row["Counter_Value"])
            name = row["Kernel_Name"]
            prev = out.get(did)
            out[did] = (name, (prev[1] if prev else 0.0) + val)
    return out


def late_pass_values(counters):
    """Per late pass: the LATE decide dispatch plus the next emit dispatch. Returns a list of summed values."""
    ids = sorted(counters)
    out = []
    for k, did in enumerate(ids):
        name, val = counters[did]
        if LATE_DECIDE.search(name):
            for did2 in ids[k + 1:]:
                n2, v2 = counters[did2]
                if EMIT.search(n2):
                    out.append((val, v2))
                    break
    return out


PYRAMID = re.compile(r"pyramid_kernel[(<]|14pyramid_kernel[EI]")  # plain or templated (block size) symbol


def frame_values(counters):
    """Per frame: the sum over the dispatches from one single-view pyramid build up to the next (the frame's
    pyramid, EARLY and LATE launches), skipping copies and fills."""
    ids = sorted(counters)
    starts = [k for k, did in enumerate(ids) if PYRAMID.search(counters[did][0])]
    out = []
    for a, b in zip(starts, starts[1:]):
        tot = 0.0
        for did in ids[a:b]:
            name, val = counters[did]
            if "rocclr" in name or "fill_u32" in name:
                continue
            tot += val
        out.append(tot)
    return out


def frame_traffic(fetch_csv: str, write_csv: str, calib_mib: float, command: str):
    fc, wc = read_counters(fetch_csv), read_counters(write_csv)
    known_kib = calib_mib * 1024.0
    f_copy = [v for n, v in fc.values() if COPY_KERNEL in n and v > 0.5 * known_kib * 0.25]
    w_copy = [v for n, v in wc.values() if COPY_KERNEL in n and v > 0.5 * known_kib * 0.25]
    if not f_copy or not w_copy:
        raise SystemExit("calibration copy kernel not found in the counter files")
    f_scale = known_kib / statistics.median(f_copy)
    w_scale = known_kib / statistics.median(w_copy)
    ff, ww = frame_values(fc), frame_values(wc)
    if not ff or not ww:
        raise SystemExit("no complete frame found")
    fetch_kib, write_kib = statistics.median(ff), statistics.median(ww)
    return {
        "unit": "C1 frame (pyramid + EARLY + LATE launches)",
        "bytes_per_launch": round((fetch_kib * f_scale + write_kib * w_scale) * 1024.0),
        "fetch_kib_raw": round(fetch_kib, 1), "write_kib_raw": round(write_kib, 1),
        "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
        "calibration": f"{COPY_KERNEL} of {calib_mib} MiB: FETCH_SIZE median {statistics.median(f_copy):.1f} KiB, "
                       f"WRITE_SIZE median {statistics.median(w_copy):.1f} KiB",
        "frames": min(len(ff), len(ww)),
        "kernel_sources": sources_digest(),
        "command": command,
    }


def traffic(fetch_csv: str, write_csv: str, calib_mib: float, command: str):
    fc, wc = read_counters(fetch_csv), read_counters(write_csv)
    known_kib = calib_mib * 1024.0
    f_copy = [v for n, v in fc.values() if COPY_KERNEL in n and v > 0.5 * known_kib * 0.25]
    w_copy = [v for n, v in wc.values() if COPY_KERNEL in n and v > 0.5 * known_kib * 0.25]
    if not f_copy or not w_copy:
        raise SystemExit("calibration copy kernel not found in the counter files")
    f_scale = known_kib / statistics.median(f_copy)
    w_scale = known_kib / statistics.median(w_copy)
    lf, lw = late_pass_values(fc), late_pass_values(wc)
    if not lf or not lw:
        raise SystemExit("no late-pass dispatches found")
    fetch_kib = statistics.median(a + b for a, b in lf)
    write_kib = statistics.median(a + b for a, b in lw)
    decide_f = statistics.median(a for a, _ in lf)
    decide_w = statistics.median(a for a, _ in lw)
    bytes_per_launch = (fetch_kib * f_scale + write_kib * w_scale) * 1024.0
    return {
        "unit": "late pass (decide_kernel LATE + emit_kernel)",
        "bytes_per_launch": round(bytes_per_launch),
        "fetch_kib_raw": round(fetch_kib, 1), "write_kib_raw": round(write_kib, 1),
        "decide_fetch_kib_raw": round(decide_f, 1), "decide_write_kib_raw": round(decide_w, 1),
        "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
        "calibration": f"{COPY_KERNEL} of {calib_mib} MiB: FETCH_SIZE median {statistics.median(f_copy):.1f} KiB, "
                       f"WRITE_SIZE median {statistics.median(w_copy):.1f} KiB",
        "late_passes": min(len(lf), len(lw)),
        "kernel_sources": sources_digest(),
        "command": command,
    }


def read_all_counters(path: str):
    """{(dispatch_id, counter): (kernel_name, value)} for every counter of a CSV counter-collection file."""
    out = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            key = (int(row["Dispatch_Id"]), row["Counter_Name"])
            prev = out.get(key)
            out[key] = (row["Kernel_Name"], (prev[1] if prev else 0.0) + float(row["Counter_Value"]))
    return out


def counters(paths, calib_mib: float, title: str, command: str) -> str:
    per = defaultdict(lambda: defaultdict(list))  # kernel -> counter -> values per dispatch
    for p in paths:
        for (_, cname), (kname, val) in read_all_counters(p).items():
            per[kname][cname].append(val)
    scale = {}
    if calib_mib:
        known_kib = calib_mib * 1024.0
        for cname in ("FETCH_SIZE", "WRITE_SIZE"):
            vals = [v for k, cs in per.items() if COPY_KERNEL in k for v in cs.get(cname, [])
                    if v > 0.5 * known_kib * 0.25]
            if vals:
                scale[cname] = known_kib / statistics.median(vals)
    names = sorted({c for cs in per.values() for c in cs})
    lines = [f"# {title}", "", f"Kernel sources: `{sources_digest()}` (blitzen_amd.build.kernel_sources_digest)", "",
             "Command:", "", "```", command, "```", "",
             "Median per dispatch." + (" FETCH_SIZE / WRITE_SIZE in KiB as read, and scaled to bytes by the "
                                       "calibration copy: " + ", ".join(f"{k} x{v:.4f}" for k, v in scale.items())
                                       if scale else ""), "",
             "| kernel | dispatches | " + " | ".join(names) + (" | fetch bytes | write bytes |" if scale else " |"),
             "|---|---|" + "---|" * len(names) + ("---|---|" if scale else "")]
    for kname, cs in sorted(per.items(), key=lambda kv: kv[0]):
        if "rocclr" in kname or "fill_u32" in kname:
            continue
        n = max(len(v) for v in cs.values())
        row = [demangle_short(kname), str(n)]
        for c in names:
            row.append(f"{statistics.median(cs[c]):.1f}" if cs.get(c) else "")
        if scale:
            for c in ("FETCH_SIZE", "WRITE_SIZE"):
                row.append(f"{statistics.median(cs[c]) * 1024 * scale[c]:.0f}" if cs.get(c) and c in scale else "")
        lines.append("| " + " | ".join(row) + " |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--title", default="Kernel trace")
    t.add_argument("--command", default="")
    t.add_argument("--out")
    r = sub.add_parser("traffic")
    r.add_argument("fetch_csv")
    r.add_argument("write_csv")
    r.add_argument("--calib-mib", type=float, required=True)
    r.add_argument("--command", default="")
    r.add_argument("--json")
    r.add_argument("--frame", action="store_true", help="per C1 frame instead of per late pass")
    k = sub.add_parser("counters")
    k.add_argument("csv", nargs="+")
    k.add_argument("--calib-mib", type=float, default=0.0)
    k.add_argument("--title", default="Counters per kernel")
    k.add_argument("--command", default="")
    k.add_argument("--out")
    a = ap.parse_args(argv)
    if a.cmd == "counters":
        text = counters(a.csv, a.calib_mib, a.title, a.command)
        if a.out:
            Path(a.out).write_text(text)
        sys.stdout.write(text)
        return
    if a.cmd == "trace":
        text = trace(a.db, a.title, a.command)
        if a.out:
            Path(a.out).write_text(text)
        sys.stdout.write(text)
    else:
        res = (frame_traffic if a.frame else traffic)(a.fetch_csv, a.write_csv, a.calib_mib, a.command)
        if a.json:
            Path(a.json).write_text(json.dumps(res, indent=1) + "\n")
        print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
