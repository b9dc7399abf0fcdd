"""Summarise rocprofv3 CSV output of the RMQ kernels into a short Markdown report.

    python -m rmq_amd.profsum --kernel k_query_xxx --queries 67108864 --bytes-per-query 195.99 \
        --stats DIR/kt_kernel_stats.csv --fetch DIR/fetch_counter_collection.csv \
        --write DIR/write_counter_collection.csv --copy-fetch DIR/copyf_counter_collection.csv \
        --copy-write DIR/copyw_counter_collection.csv [--l2 DIR/l2_counter_collection.csv] \
        [--sq DIR/sq_counter_collection.csv] [--insts DIR/insts_counter_collection.csv] \
        [--counters-json profiles/counters.json --counters-key K] \
        --command "..." > profiles/<kernel>_<workload>.md

--build-kernels k1,k2,... (with --bytes-per-build) summarises the build instead: every dispatch of
those kernels, summed and divided by the number of builds (dispatches of --build-count-kernel).

--counters-json records, under key K (bench.py's counters_key for the same kernel and workload),
the measured HBM bytes per dispatch and, with --insts (a pass over SQ_INSTS_VALU SQ_INSTS_SALU
SQ_WAVES), the wave-instructions per dispatch, so that bench.py reports roofline.traffic and the
instruction-issue roofline of the timed kernel from them.

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what the same counters read for a
plain device copy of known size (bench.py --copy-probe: 1 GiB read + 1 GiB written per launch).
"""

from __future__ import annotations

import argparse
import collections
import csv
import json
import os

from . import roofline


def _per_kernel(path, counter):
    vals = collections.defaultdict(list)
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                vals[row["Kernel_Name"]].append(float(row["Counter_Value"]))
    return vals


def _match(vals, kernel):
    hits = {k: v for k, v in vals.items() if kernel in k}
    if not hits:
        raise SystemExit(f"kernel {kernel!r} not found")
    name = max(hits, key=lambda k: len(hits[k]))
    v = hits[name]
    return name, sum(v) / len(v), len(v)


def _stats(path, kernel):
    with open(path) as f:
        for row in csv.DictReader(f):
            if kernel in row["Name"]:
                return row["Name"], int(row["Calls"]), float(row["AverageNs"]), float(row["MinNs"]), \
                    float(row["MaxNs"])
    raise SystemExit(f"kernel {kernel!r} not in stats")


def _all_stats(path):
    with open(path) as f:
        return [(row["Name"], int(row["Calls"]), float(row["TotalDurationNs"])) for row in csv.DictReader(f)]


def _sum_counter(path, counter, names):
    total = 0.0
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter and any(k in row["Kernel_Name"] for k in names):
                total += float(row["Counter_Value"])
    return total


def build_summary(a):
    """Per-build figures of the build kernels (--build-kernels, comma list of name parts), summed
    over all their dispatches and divided by the number of builds (the dispatches of
    --build-count-kernel, one per build)."""
    names = [k for k in a.build_kernels.split(",") if k]
    stats = _all_stats(a.stats)
    builds = sum(c for n, c, _ in stats if a.build_count_kernel in n)
    if builds == 0:
        raise SystemExit(f"no dispatches of {a.build_count_kernel!r}")
    rows = [(n, c, t) for n, c, t in stats if any(k in n for k in names)]
    ns = sum(t for _, _, t in rows) / builds
    fetch = _sum_counter(a.fetch, "FETCH_SIZE", names) / builds
    write = _sum_counter(a.write, "WRITE_SIZE", names) / builds
    _, cf, _ = _match(_per_kernel(a.copy_fetch, "FETCH_SIZE"), "copyBuffer")
    _, cw, _ = _match(_per_kernel(a.copy_write, "WRITE_SIZE"), "copyBuffer")
    gib_kib = float(1 << 20)
    fetch_b = fetch * 1024 * gib_kib / cf
    write_b = write * 1024 * gib_kib / cw
    algo = a.bytes_per_build
    out = [f"# {a.title}\n", "Commands (run from /tmp with TMPDIR=/tmp, repo at $R):\n"]
    out += [f"    {c}" for c in a.command]
    out += ["", "| kernel | dispatches per build | time per build |", "|---|---|---|"]
    for n, c, t in rows:
        out.append(f"| `{n}` | {c / builds:g} | {t / builds / 1e3:.2f} us |")
    out += ["", "| quantity | value |", "|---|---|",
            f"| builds traced | {builds} |",
            f"| kernel time per build | {ns / 1e3:.2f} us |",
            f"| algorithmic bytes per build (level 0 read once, every upper entry and offset written once) | {algo / 1e6:.1f} MB |",
            f"| algorithmic GB/s at kernel time | {algo / ns:.1f} ({algo / ns / roofline.HBM_PEAK_GBPS:.3f} of 8 TB/s) |",
            f"| measured traffic per build | {(fetch_b + write_b) / 1e6:.1f} MB (fetch {fetch_b / 1e6:.1f} + write {write_b / 1e6:.1f}) |",
            f"| measured GB/s at kernel time | {(fetch_b + write_b) / ns:.1f} ({(fetch_b + write_b) / ns / roofline.HBM_PEAK_GBPS:.3f} of 8 TB/s) |"]
    for line in a.note:
        out += ["", line]
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", default="")
    ap.add_argument("--queries", type=int, default=0)
    ap.add_argument("--bytes-per-query", type=float, default=0.0)
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--copy-fetch", required=True)
    ap.add_argument("--copy-write", required=True)
    ap.add_argument("--l2")
    ap.add_argument("--sq", help="counter CSV with VALUBusy, VALUUtilization, OccupancyPercent")
    ap.add_argument("--insts", help="counter CSV with SQ_INSTS_VALU, SQ_INSTS_SALU, SQ_WAVES")
    ap.add_argument("--counters-json")
    ap.add_argument("--counters-key")
    ap.add_argument("--note", action="append", default=[], help="a line printed after the table")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--title", default="")
    ap.add_argument("--build-kernels", help="build mode: comma list of kernel name parts summed per build")
    ap.add_argument("--build-count-kernel", default="k_init_words", help="one dispatch per build")
    ap.add_argument("--bytes-per-build", type=float, default=0.0)
    a = ap.parse_args()
    if a.build_kernels:
        build_summary(a)
        return

    name, calls, avg_ns, min_ns, max_ns = _stats(a.stats, a.kernel)
    _, fetch_kib, _ = _match(_per_kernel(a.fetch, "FETCH_SIZE"), a.kernel)
    _, write_kib, _ = _match(_per_kernel(a.write, "WRITE_SIZE"), a.kernel)
    _, cf, _ = _match(_per_kernel(a.copy_fetch, "FETCH_SIZE"), "copyBuffer")
    _, cw, _ = _match(_per_kernel(a.copy_write, "WRITE_SIZE"), "copyBuffer")
    gib_kib = float(1 << 20)
    f_fac, w_fac = gib_kib / cf, gib_kib / cw
    fetch_b = fetch_kib * 1024 * f_fac
    write_b = write_kib * 1024 * w_fac
    q = a.queries
    algo = a.bytes_per_query * q
    out = []
    out.append(f"# {a.title or name}\n")
    out.append("Commands (run from /tmp with TMPDIR=/tmp, repo at $R):\n")
    for c in a.command:
        out.append(f"    {c}")
    out.append("")
    out.append(f"Kernel: `{name}`\n")
    out.append("| quantity | value |")
    out.append("|---|---|")
    out.append(f"| dispatches (kernel trace) | {calls} |")
    out.append(f"| average duration | {avg_ns / 1e6:.4f} ms (min {min_ns / 1e6:.4f}, max {max_ns / 1e6:.4f}) |")
    out.append(f"| queries per dispatch | {q} |")
    out.append(f"| throughput at average duration | {q / avg_ns:.3f} Gqueries/s |")
    out.append(f"| algorithmic bytes per query (R0=32 traversal, 64-B sectors) | {a.bytes_per_query:.3f} |")
    out.append(f"| algorithmic GB/s at average duration | {algo / avg_ns:.1f} ({algo / avg_ns / 8000:.3f} of 8 TB/s) |")
    out.append(f"| FETCH_SIZE per dispatch (raw) | {fetch_kib:.1f} KiB |")
    out.append(f"| WRITE_SIZE per dispatch (raw) | {write_kib:.1f} KiB |")
    out.append(f"| copy-probe calibration (1 GiB copy) | FETCH x{f_fac:.4f}, WRITE x{w_fac:.4f} |")
    out.append(f"| measured traffic per dispatch | {(fetch_b + write_b) / 1e9:.3f} GB "
               f"(fetch {fetch_b / 1e9:.3f} + write {write_b / 1e9:.3f}) |")
    out.append(f"| measured bytes per query | {(fetch_b + write_b) / q:.1f} |")
    out.append(f"| measured L2-to-fabric GB/s | {(fetch_b + write_b) / avg_ns:.1f} |")
    if a.l2:
        _, hit, _ = _match(_per_kernel(a.l2, "TCC_HIT_sum"), a.kernel)
        _, miss, _ = _match(_per_kernel(a.l2, "TCC_MISS_sum"), a.kernel)
        out.append(f"| L2 hits / misses per dispatch | {hit:.4g} / {miss:.4g} "
                   f"(hit rate {hit / (hit + miss):.3f}, {miss / q:.2f} misses per query) |")
    if a.sq:
        for counter, label in (("VALUBusy", "VALU busy (% of GPU time)"),
                               ("VALUUtilization", "VALU utilization (% active lanes)"),
                               ("OccupancyPercent", "occupancy (% of maximum waves)")):
            vals = _per_kernel(a.sq, counter)
            if any(a.kernel in k for k in vals):
                out.append(f"| {label} | {_match(vals, a.kernel)[1]:.1f} |")
    insts = {}
    if a.insts:
        for counter in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_WAVES"):
            vals = _per_kernel(a.insts, counter)
            if any(a.kernel in k for k in vals):
                insts[counter] = _match(vals, a.kernel)[1]
        waves = insts.get("SQ_WAVES")
        if "SQ_INSTS_VALU" in insts:
            n = insts["SQ_INSTS_VALU"]
            per_wave = f", {n / waves:.0f} per wavefront" if waves else ""
            out.append(f"| VALU wave-instructions per dispatch | {n:.4g}{per_wave}, {n / q:.2f} per query |")
            out.append(f"| VALU issue at average duration | {n / avg_ns:.1f} G/s = {n / avg_ns / roofline.VALU_PEAK_FASTEST:.3f}"
                       f" of the fastest VALU op's measured issue rate ({roofline.VALU_PEAK_FASTEST} G/s, v_add_u32), "
                       f"{n / avg_ns / roofline.VALU_PEAK_FP32:.3f} of the fp32 min / select rate "
                       f"({roofline.VALU_PEAK_FP32} G/s) |")
        if "SQ_INSTS_SALU" in insts:
            n = insts["SQ_INSTS_SALU"]
            per_wave = f", {n / waves:.0f} per wavefront" if waves else ""
            out.append(f"| SALU wave-instructions per dispatch | {n:.4g}{per_wave} |")
            out.append(f"| SALU issue at average duration | {n / avg_ns:.1f} G/s = "
                       f"{n / avg_ns / roofline.SALU_PEAK_MEASURED:.3f} of the measured "
                       f"{roofline.SALU_PEAK_MEASURED} G/s |")
        if waves:
            out.append(f"| wavefronts per dispatch | {waves:.0f} ({q / waves:.1f} queries each) |")
    out.append(f"| HBM fraction from measured traffic | {(fetch_b + write_b) / avg_ns / 8000:.3f} of 8 TB/s |")
    for line in a.note:
        out.append("")
        out.append(line)
    print("\n".join(out))
    if a.counters_json and a.counters_key:
        table = {}
        if os.path.exists(a.counters_json):
            with open(a.counters_json) as f:
                table = json.load(f)
        entry = {"bytes_per_launch": round(fetch_b + write_b), "fetch_bytes": round(fetch_b),
                 "write_bytes": round(write_b), "kernel": name, "avg_ns": round(avg_ns)}
        for counter, key in (("SQ_INSTS_VALU", "valu_insts"), ("SQ_INSTS_SALU", "salu_insts"),
                             ("SQ_WAVES", "waves")):
            if counter in insts:
                entry[key] = round(insts[counter])
        table[a.counters_key] = entry
        with open(a.counters_json, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
