"""Summaries of rocprofv3 output for the registration kernels.

* kernel_stats_markdown(csv_path): the --kernel-trace --stats table
  (<dir>/<name>_kernel_stats.csv) as markdown, with readable kernel names.
* batched_stats_markdown(trace_csv): from <dir>/<name>_kernel_trace.csv,
  per kernel the dispatches of its largest grid only (the batched pipeline
  launches, apart from single-pair calls of the same kernels), with count,
  mean, min and max duration.
* traffic_summary(fetch_csv, write_csv, copy_bytes): per-kernel HBM bytes
  per launch from two --pmc runs (FETCH_SIZE, then WRITE_SIZE), each scaled
  by what the same counter reads on the streaming copy kernel, whose true
  traffic is copy_bytes read and copy_bytes written (on gfx950 one FETCH_SIZE
  unit measured 2048 bytes and one WRITE_SIZE unit 1024 bytes this way).

Usage:
  python -m phaser_amd.profiling stats <kernel_stats.csv>
  python -m phaser_amd.profiling batched <kernel_trace.csv>
  python -m phaser_amd.profiling traffic <fetch counter_collection.csv>
      <write counter_collection.csv> <copy bytes> <out.json>
"""
import csv
import json
import re
import sys
from collections import defaultdict

COPY_KERNEL = "k_copy_f4"


def short_name(name):
    """'void phz::k_fft_cols<-1, 16>(float2*, ...)' -> 'k_fft_cols<-1, 16>'."""
    n = name.strip().replace("(anonymous namespace)::", "")
    n = re.sub(r"^void\s+", "", n)
    depth = 0
    for i, ch in enumerate(n):  # cut the argument list at the top-level '('
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            n = n[:i]
            break
    return n.replace("phz::", "")


def matches(kernel, name):
    """Does the short name `name` belong to kernel label `kernel`
    ('k_fft_cols<-1>' matches every tile width of that direction)?"""
    if "<" in kernel:
        head = kernel[:-1] if kernel.endswith(">") else kernel
        return name.startswith(head + ",") or name.startswith(head + ">")
    return name == kernel or name.startswith(kernel + "<")


def kernel_stats_markdown(path):
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | avg us | min us | max us | % time |", "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.2f} |")
    return "\n".join(out)


def batched_stats(trace_csv):
    """{kernel: (grid size, [durations in us])} over each kernel's largest-grid
    dispatches in a rocprofv3 --kernel-trace CSV."""
    by = defaultdict(list)
    for r in csv.DictReader(open(trace_csv)):
        g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        by[short_name(r["Kernel_Name"])].append((g, us))
    out = {}
    for k, v in by.items():
        g = max(x[0] for x in v)
        out[k] = (g, [x[1] for x in v if x[0] == g])
    return out


def batched_stats_markdown(trace_csv):
    st = batched_stats(trace_csv)
    rows = sorted(st.items(), key=lambda kv: -sum(kv[1][1]))
    out = ["| kernel | grid (work-items) | launches | mean us | min us | max us |",
           "|---|---|---|---|---|---|"]
    for k, (g, d) in rows:
        out.append(f"| {k} | {g} | {len(d)} | {sum(d) / len(d):.1f} | {min(d):.1f} | {max(d):.1f} |")
    return "\n".join(out)


def _per_kernel(path, counter):
    """Counter values per kernel, keeping only the dispatches of the largest
    grid size (the batched pipeline launches; a profiled bench run also holds
    single-pair launches of the same kernels)."""
    rows = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name") == counter:
            rows[short_name(r["Kernel_Name"])].append((int(r["Grid_Size"]),
                                                       float(r["Counter_Value"])))
    vals = {}
    for k, v in rows.items():
        g = max(x[0] for x in v)
        vals[k] = [x[1] for x in v if x[0] == g]
    return vals


def traffic_summary(fetch_csv, write_csv, copy_bytes):
    fetch = _per_kernel(fetch_csv, "FETCH_SIZE")
    write = _per_kernel(write_csv, "WRITE_SIZE")

    def mean(v):
        return sum(v) / len(v)

    copy_f = [mean(v) for k, v in fetch.items() if matches(COPY_KERNEL, k)]
    copy_w = [mean(v) for k, v in write.items() if matches(COPY_KERNEL, k)]
    if not copy_f or not copy_w:
        raise RuntimeError("the copy kernel is missing from the counter runs")
    scale_f = copy_bytes / copy_f[0]  # bytes per counter unit
    scale_w = copy_bytes / copy_w[0]
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        f, w = mean(fetch[k]) * scale_f, mean(write[k]) * scale_w
        kernels[k] = {"fetch_bytes_per_launch": int(f), "write_bytes_per_launch": int(w),
                      "hbm_bytes_per_launch": int(f + w), "launches": len(fetch[k])}
    return {"copy_bytes": copy_bytes, "fetch_scale_bytes_per_unit": scale_f,
            "write_scale_bytes_per_unit": scale_w, "kernels": kernels}


def lookup(summary, kernel):
    """Mean over every entry of `summary['kernels']` that matches the kernel
    label (e.g. both tile widths of k_fft_cols<-1>)."""
    hits = [v for k, v in summary.get("kernels", {}).items() if matches(kernel, k)]
    if not hits:
        return None
    return int(sum(h["hbm_bytes_per_launch"] * h["launches"] for h in hits)
               / sum(h["launches"] for h in hits))


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        print(kernel_stats_markdown(sys.argv[2]))
    elif sys.argv[1] == "batched":
        print(batched_stats_markdown(sys.argv[2]))
    elif sys.argv[1] == "traffic":
        s = traffic_summary(sys.argv[2], sys.argv[3], int(sys.argv[4]))
        with open(sys.argv[5], "w") as f:
            json.dump(s, f, indent=1)
        print(json.dumps(s, indent=1))
    else:
        raise SystemExit(__doc__)
