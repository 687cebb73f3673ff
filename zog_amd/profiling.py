"""Summaries of rocprofv3 CSV output for the measurement files.

  python -m zog_amd.profiling stats <run_kernel_stats.csv>
      markdown table of the kernel statistics (per kernel: calls, total ms,
      average / min / max microseconds, share of device time)

  python -m zog_amd.profiling dispatches <run_kernel_trace.csv> [--match s] [--last N]
      per-dispatch duration percentiles of the kernels whose name has s

  python -m zog_amd.profiling reduce <run_counter_collection.csv> <out.csv>
      the counter values summed per (kernel, counter), with the number of
      dispatches: a run of a few hundred thousand dispatches shrinks to a few
      rows, and `traffic` reads either form

  python -m zog_amd.profiling traffic --fetch F.csv --write W.csv \
      --copy-fetch CF.csv --copy-write CW.csv --copy-bytes 1073741824 \
      [--units KERNEL=COUNT ...] -o out.json
      HBM bytes per dispatch of every kernel from one FETCH_SIZE and one
      WRITE_SIZE counter-collection run (KiB per dispatch), each scaled by
      what the same counter reads for a plain copy of known size (the copy
      kernel's mean dispatch in the calibration run, whose streaming sum of
      the same size checks the read factor); the Brandes
      forward pull and push kernels are also summed under
      "bc_forward_kernel+bc_push_kernel", as bench.py reports them together.
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict


def _short(name: str) -> str:
    """Kernel name without its argument list (template arguments kept)."""
    name = name.replace("(anonymous namespace)::", "")
    base = name.split("(")[0].strip()
    if base.startswith("void "):
        base = base[5:]
    return base


def kernel_stats_markdown(path: str) -> str:
    rows = list(csv.DictReader(open(path)))
    out = ["| Kernel | Calls | Total ms | Avg us | Min us | Max us | % |", "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        out.append("| `%s` | %s | %.2f | %.1f | %.1f | %.1f | %.2f |" % (
            r["Name"][:70], r["Calls"], float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3,
            float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    return "\n".join(out)


def _counter_per_kernel(path: str, counter: str):
    """{kernel: (sum of counter over dispatches, dispatches)} and the
    largest single dispatch value."""
    tot = defaultdict(float)
    cnt = defaultdict(int)
    biggest = 0.0
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        v = float(r["Counter_Value"])
        k = _short(r["Kernel_Name"])
        n = int(r.get("Dispatches") or 1)  # reduced files: sums over n dispatches
        tot[k] += v
        cnt[k] += n
        biggest = max(biggest, v / n)
    return tot, cnt, biggest


def reduce_counters(path: str, out: str) -> None:
    tot = defaultdict(float)
    cnt = defaultdict(int)
    for r in csv.DictReader(open(path)):
        key = (r["Kernel_Name"], r["Counter_Name"])
        tot[key] += float(r["Counter_Value"])
        cnt[key] += int(r.get("Dispatches") or 1)
    with open(out, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value", "Dispatches"])
        for (k, c), v in sorted(tot.items()):
            w.writerow([k, c, repr(v), cnt[(k, c)]])


def traffic(fetch, write, copy_fetch, copy_write, copy_bytes, units=None):
    ft, fc, _ = _counter_per_kernel(fetch, "FETCH_SIZE")
    wt, wc, _ = _counter_per_kernel(write, "WRITE_SIZE")
    cft, cfc, _ = _counter_per_kernel(copy_fetch, "FETCH_SIZE")
    _, _, cw = _counter_per_kernel(copy_write, "WRITE_SIZE")
    # the calibration run holds a copy kernel (reads and writes copy_bytes)
    # and a streaming sum (reads copy_bytes, writes a few bytes); counters are
    # in KiB.  The copy sets the factors, the sum checks the read factor on a
    # kernel that only reads.
    copy_k = [k for k in cft if "reduce" not in k.lower()]
    sum_k = [k for k in cft if "reduce" in k.lower()]
    cf = max(cft[k] / cfc[k] for k in copy_k)
    f_scale = copy_bytes / (cf * 1024.0)
    w_scale = copy_bytes / (cw * 1024.0)
    out = {"_calibration": {"copy_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
                            "fetch_factor": f_scale, "write_factor": w_scale}}
    if sum_k:
        sf = max(cft[k] / cfc[k] for k in sum_k)
        out["_calibration"]["sum_fetch_kib"] = sf
        out["_calibration"]["sum_read_bytes_scaled"] = sf * 1024.0 * f_scale
        out["_calibration"]["sum_check"] = sf * 1024.0 * f_scale / copy_bytes
    for k in sorted(set(ft) | set(wt)):
        n = max(fc.get(k, 0), wc.get(k, 0))
        if not n:
            continue
        fb = ft.get(k, 0.0) * 1024.0 * f_scale / max(fc.get(k, 1), 1)
        wb = wt.get(k, 0.0) * 1024.0 * w_scale / max(wc.get(k, 1), 1)
        out[k] = {"dispatches": n, "fetch_bytes_per_dispatch": fb, "write_bytes_per_dispatch": wb,
                  "hbm_bytes_per_dispatch": fb + wb}
    a, b = "bc_forward_kernel", "bc_push_kernel"
    if a in out and b in out:
        n = out[a]["dispatches"] + out[b]["dispatches"]
        tot = out[a]["hbm_bytes_per_dispatch"] * out[a]["dispatches"] + \
            out[b]["hbm_bytes_per_dispatch"] * out[b]["dispatches"]
        out[a + "+" + b] = {"dispatches": n, "hbm_bytes_per_dispatch": tot / n}
    # units of work a kernel's dispatches served in the profiled run, when
    # that is not one per dispatch (a PageRank run launches its iterations,
    # one closing launch and launches that return at once): the bytes per
    # unit are the kernel's total over them
    for k, u in (units or {}).items():
        if k in out and u > 0:
            out[k]["units"] = u
            out[k]["hbm_bytes_per_unit"] = out[k]["hbm_bytes_per_dispatch"] * out[k]["dispatches"] / u
    return out


def dispatch_markdown(path: str, match: str, small_us: float = 20.0, last: int = 0) -> str:
    """Per-dispatch duration distribution of the kernels whose name contains
    `match`, from a rocprofv3 kernel trace: dispatches, total ms, mean, 10th /
    50th / 90th percentile and largest microseconds, and the dispatches
    shorter than `small_us` with their total.  With `last`, only each
    kernel's last `last` dispatches (in start order) count."""
    durs = defaultdict(list)
    for r in csv.DictReader(open(path)):
        name = _short(r["Kernel_Name"])
        if match in name:
            durs[name].append((int(r["Start_Timestamp"]),
                               (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0))
    out = ["| Kernel | Dispatches | Total ms | Mean us | p10 us | p50 us | p90 us | Max us | Under %g us | Their ms |"
           % small_us, "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    rows = []
    for name, tv in durs.items():
        tv.sort()
        v = [d for _, d in (tv[-last:] if last else tv)]
        rows.append((name, v))
    for name, v in sorted(rows, key=lambda kv: -sum(kv[1])):
        v = sorted(v)
        small = [x for x in v if x < small_us]
        out.append("| `%s` | %d | %.1f | %.1f | %.1f | %.1f | %.1f | %.1f | %d | %.1f |" % (
            name, len(v), sum(v) / 1000.0, sum(v) / len(v), v[len(v) // 10], v[len(v) // 2],
            v[(9 * len(v)) // 10], v[-1], len(small), sum(small) / 1000.0))
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    d = sub.add_parser("dispatches")
    d.add_argument("csv")
    d.add_argument("--match", default="")
    d.add_argument("--last", type=int, default=0, help="only each kernel's last N dispatches")
    r = sub.add_parser("reduce")
    r.add_argument("csv")
    r.add_argument("out")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-bytes", type=int, default=1 << 30)
    t.add_argument("--units", action="append", default=[],
                   help="KERNEL=COUNT: units of work (iterations) the kernel's dispatches served")
    t.add_argument("-o", "--out", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(kernel_stats_markdown(a.csv))
    elif a.cmd == "dispatches":
        print(dispatch_markdown(a.csv, a.match, last=a.last))
    elif a.cmd == "reduce":
        reduce_counters(a.csv, a.out)
    else:
        units = {}
        for spec in a.units:
            k, _, c = spec.rpartition("=")
            units[k] = int(c)
        res = traffic(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes, units)
        with open(a.out, "w") as fh:
            json.dump(res, fh, indent=1)
        json.dump(res, sys.stdout, indent=1)


if __name__ == "__main__":
    main()
