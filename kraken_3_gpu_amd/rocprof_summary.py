"""Summarise rocprofv3 CSV output of a bench.py run into the short text kept
under profiles/ and the per-launch HBM traffic bench.py reports.

HBM traffic of a kernel launch comes from two counter passes (FETCH_SIZE, then
WRITE_SIZE, KiB per dispatch). Each is scaled by what the same counter reads
for a device copy of known size (bench.py --copy-probe): on gfx950 a 256 MiB
copy reads back FETCH_SIZE = 128 MiB and WRITE_SIZE = 256 MiB, so the fetch
counter is scaled by 2.0 and the write counter by 1.0.

  python -m kraken_3_gpu_amd.rocprof_summary --kernel kx_k_persist \\
      --stats DIR/kt_kernel_stats.csv --trace DIR/kt_kernel_trace.csv \\
      --fetch F.csv --write W.csv --copy-fetch CF.csv --copy-write CW.csv \\
      [--sq S1.csv S2.csv] --workload "..." --creatures N --steps-per-launch K --command "..." \\
      --out profiles/NAME.md --json profiles/NAME_traffic.json
"""
import argparse
import collections
import csv
import json


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def counter_per_dispatch(path, kernel, counter):
    """{dispatch id: summed counter value} for dispatches of `kernel`."""
    out = collections.defaultdict(float)
    for r in _rows(path):
        if kernel in r["Kernel_Name"] and r["Counter_Name"] == counter:
            out[int(r["Dispatch_Id"])] += float(r["Counter_Value"])
    return dict(out)


def copy_factor(path, counter, bytes_moved):
    """Bytes per counter KiB, from the largest copyBuffer dispatch of a copy probe."""
    vals = counter_per_dispatch(path, "copyBuffer", counter)
    if not vals:
        return None
    kib = max(vals.values())
    return bytes_moved / (kib * 1024.0) if kib > 0 else None


def trace_durations(path, kernel, min_ms=0.0):
    """Durations (ms) of the kernel's dispatches in issue order, those shorter
    than min_ms left out (launches of fewer steps)."""
    out = []
    for r in _rows(path):
        if kernel in r["Kernel_Name"]:
            out.append((int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6))
    return [d for _, d in sorted(out) if d >= min_ms]


def first_dispatch_counters(path, kernel):
    """{counter: value} of the kernel's first dispatch in one counter run."""
    rows = [r for r in _rows(path) if kernel in r["Kernel_Name"]]
    if not rows:
        return {}
    first = min(int(r["Dispatch_Id"]) for r in rows)
    out = {}
    for r in rows:
        if int(r["Dispatch_Id"]) == first:
            out[r["Counter_Name"]] = out.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--stats", required=True)
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=int, default=256 << 20)
    ap.add_argument("--sq", nargs="*", default=[])
    ap.add_argument("--workload", required=True)
    ap.add_argument("--creatures", type=int, required=True)
    ap.add_argument("--steps-per-launch", type=int, required=True)
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--out", required=True)
    ap.add_argument("--json")
    ap.add_argument("--note", action="append", default=[], help="free text lines appended to the summary")
    ap.add_argument("--min-ms", type=float, default=0.0,
                    help="trace statistics over dispatches at least this long (the full-length launches)")
    ap.add_argument("--last", type=int, default=0,
                    help="also report the mean of the last N of those dispatches (a bench leg's launches)")
    a = ap.parse_args()

    stats = _rows(a.stats)
    traffic = bool(a.fetch and a.write and a.copy_fetch and a.copy_write)
    fetch = counter_per_dispatch(a.fetch, a.kernel, "FETCH_SIZE") if traffic else {}
    write = counter_per_dispatch(a.write, a.kernel, "WRITE_SIZE") if traffic else {}
    ff = copy_factor(a.copy_fetch, "FETCH_SIZE", a.copy_bytes) if traffic else None
    fw = copy_factor(a.copy_write, "WRITE_SIZE", a.copy_bytes) if traffic else None
    # the kernel's first dispatch in each run: the same launch in every pass
    fetch_kib = fetch[min(fetch)] if fetch else None
    write_kib = write[min(write)] if write else None
    fetch_b = fetch_kib * 1024 * ff if fetch_kib is not None and ff else None
    write_b = write_kib * 1024 * fw if write_kib is not None and fw else None

    lines = [f"# {a.kernel}: {a.workload}", "", "Commands (run from the repository root, `cd /tmp && export TMPDIR=/tmp` first):", ""]
    lines += [f"    {c}" for c in a.command]
    lines += ["", "## Kernel trace (rocprofv3 --kernel-trace --stats)", "",
              "| kernel | calls | average ms | share % |", "|---|---|---|---|"]
    for r in stats[:12]:
        lines.append(f"| `{r['Name'][:70]}` | {r['Calls']} | {float(r['AverageNs']) * 1e-6:.3f} | {float(r['Percentage']):.2f} |")
    if a.trace:
        d = trace_durations(a.trace, a.kernel, a.min_ms)
        if d:
            s = sorted(d)
            sel = f" of at least {a.min_ms} ms" if a.min_ms else ""
            lines += ["", f"{a.kernel} dispatches{sel}: {len(d)}, mean {sum(d) / len(d):.3f} ms, median "
                          f"{s[len(s) // 2]:.3f} ms, min {s[0]:.3f} ms, max {s[-1]:.3f} ms"]
            if a.last and len(d) >= a.last:
                t = d[-a.last:]
                lines += [f"last {a.last} of them: mean {sum(t) / len(t):.3f} ms"]
            if len(d) <= 12:
                lines += [f"{a.kernel} dispatch durations (ms): " + ", ".join(f"{x:.3f}" for x in d)]
    if traffic:
        lines += ["", "## HBM traffic per launch (rocprofv3 --pmc, one counter per run)", "",
                  "| counter | KiB per dispatch | copy-probe scale | bytes per launch |", "|---|---|---|---|",
                  f"| FETCH_SIZE | {fetch_kib} | {ff} | {fetch_b} |",
                  f"| WRITE_SIZE | {write_kib} | {fw} | {write_b} |"]
    sq = {}
    for p in a.sq:
        sq.update(first_dispatch_counters(p, a.kernel))
    if sq:
        waves = sq.get("SQ_WAVES", 0.0)
        wc = sq.get("SQ_WAVE_CYCLES", 0.0)
        lines += ["", "## SQ counters (the kernel's first dispatch of each counter run)", "", "| counter | total | per wave | / SQ_WAVE_CYCLES |", "|---|---|---|---|"]
        for k in sorted(sq):
            v = sq[k]
            lines.append(f"| {k} | {v:.4e} | {v / waves if waves else 0:.4e} | {v / wc if wc else 0:.4f} |")
        if sq.get("SQ_INSTS_VALU") and "SQ_THREAD_CYCLES_VALU" in sq:
            lines.append("")
            lines.append(f"Active lanes per VALU instruction: {sq['SQ_THREAD_CYCLES_VALU'] / (sq['SQ_INSTS_VALU'] * 64):.3f}")
    if a.note:
        lines += [""] + a.note
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    if not a.json:
        return
    with open(a.json, "w") as f:
        json.dump({"kernel": a.kernel, "workload": a.workload, "creatures": a.creatures,
                   "steps_per_launch": a.steps_per_launch, "fetch_kib": fetch_kib, "write_kib": write_kib,
                   "fetch_scale": ff, "write_scale": fw, "fetch_bytes": fetch_b, "write_bytes": write_b,
                   "traffic_bytes": (fetch_b + write_b) if fetch_b is not None and write_b is not None else None,
                   "commands": a.command}, f, indent=1)


if __name__ == "__main__":
    main()
