"""Summarize a rocprofv3 CSV run per kernel, restricted to a roctx range.

Usage::

    rocprofv3 --kernel-trace --marker-trace --stats --output-format csv \
        -d OUT -o run -- python bench.py ...
    python -m impuritymodel_amd.profile_summary OUT/run fit_timed_surrogate

prints a Markdown table (dispatches, total, average, min, max in ms) of the
kernels that started inside every instance of the named roctx range
(``bench.py`` opens one around each timed region), followed by rocprofv3's
own whole-run ``kernel_stats`` lines for the same kernels.
"""

import csv
import sys
from collections import defaultdict


def _short(name, width=90):
    name = name.replace("imped::(anonymous namespace)::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def summarize(prefix, range_name, top=12):
    ranges = []
    with open(prefix + "_marker_api_trace.csv") as f:
        for row in csv.DictReader(f):
            if row["Function"] == range_name:
                ranges.append((int(row["Start_Timestamp"]), int(row["End_Timestamp"])))
    if not ranges:
        raise SystemExit("no roctx range named %r in %s_marker_api_trace.csv" % (range_name, prefix))
    per = defaultdict(list)
    with open(prefix + "_kernel_trace.csv") as f:
        for row in csv.DictReader(f):
            t0, t1 = int(row["Start_Timestamp"]), int(row["End_Timestamp"])
            if any(a <= t0 and t1 <= b for a, b in ranges):
                per[row["Kernel_Name"]].append((t1 - t0) * 1e-6)
    total = sum(sum(v) for v in per.values())
    lines = ["Kernels inside roctx range `%s` (%d instance(s), %.3f ms of kernel time):" % (
        range_name, len(ranges), total), "",
        "| kernel | dispatches | total ms | avg ms | min ms | max ms | share |", "|---|---|---|---|---|---|---|"]
    for name, d in sorted(per.items(), key=lambda kv: -sum(kv[1]))[:top]:
        lines.append("| `%s` | %d | %.4f | %.4f | %.4f | %.4f | %.1f%% |" % (
            _short(name), len(d), sum(d), sum(d) / len(d), min(d), max(d), 100 * sum(d) / total))
    lines += ["", "Whole-run rocprofv3 kernel_stats (all dispatches, including set-up) for the same kernels:", "",
              "| kernel | calls | avg ms | total ms |", "|---|---|---|---|"]
    keep = set(sorted(per, key=lambda k: -sum(per[k]))[:top])
    with open(prefix + "_kernel_stats.csv") as f:
        for row in csv.DictReader(f):
            if row["Name"] in keep:
                lines.append("| `%s` | %s | %.4f | %.3f |" % (_short(row["Name"]), row["Calls"],
                                                            float(row["AverageNs"]) * 1e-6,
                                                            float(row["TotalDurationNs"]) * 1e-6))
    return "\n".join(lines)


def pmc_traffic(counter_csv, kernel, copy_kernel, copy_bytes, grids=None):
    """Per-dispatch counter averages of ``kernel`` scaled to bytes.

    ``counter_csv`` is a rocprofv3 ``--pmc`` counter_collection.csv (one
    counter per run: FETCH_SIZE or WRITE_SIZE, reported in KiB).  The copy
    kernel of known size gives the factor bytes-per-counter-unit; the
    kernel's average counter value times that factor is its traffic.
    ``grids``: only dispatches of ``kernel`` with these grid sizes (work
    items) count, e.g. one batch size of a kernel that also runs at others.
    """
    vals = defaultdict(list)
    counter = None
    with open(counter_csv) as f:
        for row in csv.DictReader(f):
            counter = row["Counter_Name"]
            name = row["Kernel_Name"]
            key = "copy" if copy_kernel in name else ("kernel" if kernel in name else None)
            if key == "kernel" and grids is not None and int(row["Grid_Size"]) not in grids:
                key = None
            if key:
                vals[(key, row["Dispatch_Id"])].append(float(row["Counter_Value"]))
    per = defaultdict(list)
    for (key, _), v in vals.items():
        per[key].append(sum(v))  # sum over XCDs / instances of one dispatch
    if not per["copy"] or not per["kernel"]:
        raise SystemExit("no dispatches of %r or %r in %s" % (kernel, copy_kernel, counter_csv))
    copy_units = max(per["copy"])
    factor = copy_bytes / copy_units
    avg = sum(per["kernel"]) / len(per["kernel"])
    return dict(counter=counter, copy_bytes=copy_bytes, copy_counter=copy_units, bytes_per_unit=factor,
                dispatches=len(per["kernel"]), counter_avg=avg, bytes_avg=avg * factor)


if __name__ == "__main__":
    if sys.argv[1] == "--pmc":
        grids = {int(g) for g in sys.argv[6].split(",")} if len(sys.argv) > 6 else None
        out = pmc_traffic(sys.argv[2], sys.argv[3], sys.argv[4], float(sys.argv[5]), grids)
        print(" ".join("%s=%s" % kv for kv in out.items()))
    else:
        print(summarize(sys.argv[1], sys.argv[2]))
