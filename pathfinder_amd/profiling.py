"""Summaries of rocprofv3 CSV output for the renderer's kernels.

    python -m pathfinder_amd.profiling stats DIR [--frames K] [--command CMD]
    python -m pathfinder_amd.profiling traffic FETCH_DIR WRITE_DIR --copy-bytes N [--command CMD]
    python -m pathfinder_amd.profiling counters DIR [--command CMD]

`stats` turns a `rocprofv3 --kernel-trace --stats --output-format csv` run
into a markdown table of per-kernel time. `traffic` turns two
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of
`bench.py --copy-calibration` into HBM bytes per kernel for the last frame of
the run: each counter (KiB per dispatch) is scaled by the factor that makes
the same counter read N bytes for the runtime's own copy of N bytes (the
last three copy dispatches of the run). Both print markdown; the traffic table
has one row per kernel that bench.py reads back (`read_traffic`). `counters`
sums every counter of a `--pmc` run per kernel over the last frame.
"""
import argparse
import csv
import glob
import os
import re
import sys

_KERNEL_RE = re.compile(r"(k_\w+)(?:<[^>]*>)?\(")
COPY_KERNEL = "__amd_rocclr_copyBuffer"
FRAME_MARK = "k_frame_reset"


def short_name(full):
    m = _KERNEL_RE.search(full)
    return m.group(1) if m else full


def _one(pattern):
    hits = sorted(glob.glob(pattern, recursive=True))
    if not hits:
        raise FileNotFoundError(pattern)
    return hits[-1]


def kernel_stats(directory):
    rows = list(csv.DictReader(open(_one(os.path.join(directory, "**", "*kernel_stats.csv")))))
    out = []
    for r in rows:
        out.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                    "total_ms": float(r["TotalDurationNs"]) / 1e6,
                    "avg_us": float(r["AverageNs"]) / 1e3, "pct": float(r["Percentage"])})
    return out


def frame_stats(directory, frames):
    """Per-kernel launches and duration over the last `frames` frames of a
    kernel trace (from the frames-th last frame-reset dispatch on): the timed
    frames of a bench.py run, which come last."""
    path = _one(os.path.join(directory, "**", "*kernel_trace.csv"))
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    marks = [i for i, r in enumerate(rows) if short_name(r["Kernel_Name"]) == FRAME_MARK]
    if len(marks) < frames:
        raise ValueError("fewer than %d frames in the trace" % frames)
    per = {}
    for r in rows[marks[-frames]:]:
        k = short_name(r["Kernel_Name"])
        e = per.setdefault(k, [0, 0.0])
        e[0] += 1
        e[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    return per


def stats_markdown(directory, command=None, frames=0):
    lines = ["# Kernel time, rocprofv3 --kernel-trace --stats", ""]
    if command:
        lines += ["Command (from the repository root):", "", "    " + command, ""]
    if frames:
        per = frame_stats(directory, frames)
        lines += ["The last %d frames of the run (the timed frames of bench.py), from the kernel "
                  "trace; us per frame and per launch:" % frames, "",
                  "| kernel | launches | us per frame | average us per launch |",
                  "|---|---:|---:|---:|"]
        for k in sorted(per, key=lambda k: -per[k][1]):
            n, us = per[k]
            lines.append("| %s | %d | %.1f | %.1f |" % (k, n, us / frames, us / n))
        lines += ["", "The whole run (autotune, warm-up and timed frames), rocprofv3 --stats:", ""]
    lines += ["| kernel | calls | total ms | average us | % |", "|---|---:|---:|---:|---:|"]
    for r in kernel_stats(directory):
        lines.append("| %s | %d | %.3f | %.1f | %.2f |" % (r["kernel"], r["calls"], r["total_ms"],
                                                          r["avg_us"], r["pct"]))
    return "\n".join(lines) + "\n"


def _dispatches(directory, counter):
    path = _one(os.path.join(directory, "**", "*counter_collection.csv"))
    rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [(short_name(r["Kernel_Name"]), float(r["Counter_Value"])) for r in rows]


def _last_frame(disp):
    """Dispatches from the last frame-reset kernel up to the calibration copies."""
    marks = [i for i, (k, _) in enumerate(disp) if k == FRAME_MARK]
    if not marks:
        raise ValueError("no %s dispatch: not a bench.py run" % FRAME_MARK)
    frame = disp[marks[-1]:]
    return [(k, v) for k, v in frame if k.startswith("k_")]


def _copy_factor(disp, nbytes):
    copies = [v for k, v in disp if k == COPY_KERNEL]
    if len(copies) < 3:
        raise ValueError("no calibration copies (run bench.py --copy-calibration)")
    kib = sum(copies[-3:]) / 3.0
    return nbytes / (kib * 1024.0)


def traffic(fetch_dir, write_dir, copy_bytes):
    fd, wd = _dispatches(fetch_dir, "FETCH_SIZE"), _dispatches(write_dir, "WRITE_SIZE")
    ff, fw = _copy_factor(fd, copy_bytes), _copy_factor(wd, copy_bytes)
    per = {}
    for disp, key, factor in ((_last_frame(fd), "fetch", ff), (_last_frame(wd), "write", fw)):
        for k, v in disp:
            e = per.setdefault(k, {"launches_fetch": 0, "launches_write": 0,
                                   "fetch": 0.0, "write": 0.0})
            e[key] += v * 1024.0 * factor
            e["launches_" + key] += 1
    return per, ff, fw


def traffic_markdown(fetch_dir, write_dir, copy_bytes, command=None):
    per, ff, fw = traffic(fetch_dir, write_dir, copy_bytes)
    lines = ["# HBM traffic per kernel, one frame, rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE", ""]
    if command:
        lines += ["Commands (from the repository root):", ""]
        lines += ["    " + c for c in command.split(" ;; ")] + [""]
    lines += ["Scale: a device copy of %d bytes reads FETCH_SIZE x %.4f and WRITE_SIZE x %.4f; "
              "every figure below is the counter (KiB) x 1024 x that factor, summed over the "
              "kernel's dispatches in the last frame of the run." % (copy_bytes, ff, fw), ""]
    lines += ["| kernel | launches | fetch bytes | write bytes | traffic bytes |",
              "|---|---:|---:|---:|---:|"]
    for k in sorted(per, key=lambda k: -(per[k]["fetch"] + per[k]["write"])):
        e = per[k]
        lines.append("| %s | %d | %d | %d | %d |" % (k, e["launches_fetch"], round(e["fetch"]),
                                                    round(e["write"]),
                                                    round(e["fetch"] + e["write"])))
    return "\n".join(lines) + "\n"


def counters_markdown(directory, command=None):
    path = _one(os.path.join(directory, "**", "*counter_collection.csv"))
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    names = [short_name(r["Kernel_Name"]) for r in rows]
    marks = [i for i, k in enumerate(names) if k == FRAME_MARK]
    if not marks:
        raise ValueError("no %s dispatch: not a bench.py run" % FRAME_MARK)
    per, counters = {}, []
    for k, r in zip(names[marks[-1]:], rows[marks[-1]:]):
        if not k.startswith("k_"):
            continue
        c = r["Counter_Name"]
        if c not in counters:
            counters.append(c)
        per.setdefault(k, {}).setdefault(c, 0.0)
        per[k][c] += float(r["Counter_Value"])
    lines = ["# Counters per kernel, one frame, rocprofv3 --pmc", ""]
    if command:
        lines += ["Command (from the repository root):", "", "    " + command, ""]
    lines += ["| kernel | " + " | ".join(counters) + " |",
              "|---|" + "---:|" * len(counters)]
    for k in sorted(per, key=lambda k: -per[k].get(counters[0], 0.0)):
        lines.append("| %s | " % k + " | ".join("%.4g" % per[k].get(c, 0.0) for c in counters)
                     + " |")
    return "\n".join(lines) + "\n"


def read_traffic(path):
    """{kernel: (launches, fetch, write, traffic)} from a traffic table; {} if absent."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].startswith("k_"):
            try:
                out[cells[0]] = tuple(int(c) for c in cells[1:])
            except ValueError:
                pass
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pathfinder_amd.profiling")
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command")
    s.add_argument("--frames", type=int, default=0)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--copy-bytes", type=int, default=256 << 20)
    t.add_argument("--command")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--command")
    args = ap.parse_args(argv)
    if args.what == "stats":
        sys.stdout.write(stats_markdown(args.dir, args.command, args.frames))
    elif args.what == "counters":
        sys.stdout.write(counters_markdown(args.dir, args.command))
    else:
        sys.stdout.write(traffic_markdown(args.fetch_dir, args.write_dir, args.copy_bytes,
                                          args.command))


if __name__ == "__main__":
    main()
