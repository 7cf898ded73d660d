"""Summaries of rocprofv3 CSV output, as the tables kept under profiles/.

    python3 -m thievory_amd.profsum stats  PREFIX_kernel_stats.csv [--top N]
    python3 -m thievory_amd.profsum trace  PREFIX_kernel_trace.csv [--match k_expand]
    python3 -m thievory_amd.profsum pmc    PREFIX_counter_collection.csv [--factor F]

stats : the per-kernel table of `rocprofv3 --stats` (calls, total, average).
trace : per-kernel calls / total / average from a kernel trace, keeping
        template arguments apart (k_expand<0,...> and k_expand<5,...> are
        different kernels).
pmc   : per-kernel sum and per-dispatch average of each counter (rocprofv3
        --pmc reports memory counters in KiB per dispatch), divided by the
        calibration factor F (the counter's reading for a copy of known size
        over the ideal reading).
"""
import argparse
import collections
import csv
import sys


def _short(name, width=72):
    name = name.replace("void ", "").replace("thv::", "")
    return name if len(name) <= width else name[:width - 3] + "..."


def stats(path, top):
    rows = list(csv.DictReader(open(path)))
    total = sum(float(r["TotalDurationNs"]) for r in rows) or 1.0
    print("| kernel | calls | total ms | avg ms | share |")
    print("|---|---:|---:|---:|---:|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        t = float(r["TotalDurationNs"])
        print(f"| `{_short(r['Name'])}` | {r['Calls']} | {t / 1e6:.3f} | "
              f"{float(r['AverageNs']) / 1e6:.4f} | {100 * t / total:.1f}% |")


def trace(path, match):
    agg = collections.OrderedDict()
    for r in csv.DictReader(open(path)):
        n = r["Kernel_Name"]
        if match and match not in n:
            continue
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        a = agg.setdefault(n, [0, 0])
        a[0] += 1
        a[1] += d
    print("| kernel | calls | total ms | avg ms |")
    print("|---|---:|---:|---:|")
    for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{_short(n, 96)}` | {c} | {t / 1e6:.3f} | {t / c / 1e6:.4f} |")


def pmc(path, factor):
    agg = collections.OrderedDict()
    for r in csv.DictReader(open(path)):
        key = (r["Kernel_Name"], r["Counter_Name"])
        a = agg.setdefault(key, [set(), 0.0])
        a[0].add(r.get("Dispatch_Id") or r.get("Correlation_Id"))
        a[1] += float(r["Counter_Value"])
    print(f"| kernel | counter | dispatches | sum | per dispatch | per dispatch / {factor} |")
    print("|---|---|---:|---:|---:|---:|")
    for (n, c), (ds, s) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        k = len(ds)
        print(f"| `{_short(n, 96)}` | {c} | {k} | {s:.0f} | {s / k:.1f} | "
              f"{s / k / factor:.1f} |")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "trace", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--top", type=int, default=12)
    ap.add_argument("--match", default="")
    ap.add_argument("--factor", type=float, default=1.0)
    a = ap.parse_args(argv)
    if a.mode == "stats":
        stats(a.csv, a.top)
    elif a.mode == "trace":
        trace(a.csv, a.match)
    else:
        pmc(a.csv, a.factor)


if __name__ == "__main__":
    sys.exit(main())
