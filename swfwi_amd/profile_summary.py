"""Summaries of rocprofv3 CSV output for the profiles/ folder.

    python -m swfwi_amd.profile_summary trace <dir>/<prefix>_kernel_trace.csv [--kernels REGEX]
        per kernel variant (template arguments, blocks) whose name matches
        REGEX (default: the step kernels): calls and mean / median / min
        duration;
    python -m swfwi_amd.profile_summary stats <dir>/<prefix>_kernel_stats.csv [--top N]
        the --stats summary as a table (calls, total ms, average us, %);
    python -m swfwi_amd.profile_summary hbm <copy_fetch.csv> <copy_write.csv> \
            <fetch.csv> <write.csv> --copy-floats N [--select VARIANT:BLOCKS ...]
        HBM traffic per dispatch from FETCH_SIZE / WRITE_SIZE (KiB), scaled by
        what the same counters read for a plain device copy of N floats, and
        the dispatch-weighted average over the selected variants (the launches
        of one workload), printed as a `traffic_bytes_per_launch` line.
"""
import argparse
import collections
import csv
import re
import statistics
import sys


def variant(name):
    m = re.search(r"(k_\w+)(<[^>]*>)?", name)
    return (m.group(1) + (m.group(2) or "")) if m else name.split("(")[0]


def trace(path, out, kernels=r"k_(fwd|bwd)_"):
    rows = list(csv.DictReader(open(path)))
    d = collections.defaultdict(list)
    for r in rows:
        if not re.search(kernels, r["Kernel_Name"]):
            continue
        blocks = int(r["Grid_Size_X"]) // int(r["Workgroup_Size_X"])
        d[(variant(r["Kernel_Name"]), blocks)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out.write("| kernel | blocks | calls | mean us | median us | min us |\n|---|---|---|---|---|---|\n")
    for (k, b), v in sorted(d.items()):
        out.write(f"| `{k}` | {b} | {len(v)} | {statistics.mean(v):.2f} | {statistics.median(v):.2f} | {min(v):.2f} |\n")


def stats(path, out, top=16):
    rows = list(csv.DictReader(open(path)))
    out.write("| kernel | calls | total ms | average us | % |\n|---|---|---|---|---|\n")
    for r in rows[:top]:
        out.write(f"| `{variant(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                  f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |\n")


def counter(path, pick):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if pick(r["Kernel_Name"]):
            blocks = int(r["Grid_Size"]) // int(r["Workgroup_Size"])
            d[(variant(r["Kernel_Name"]), blocks)].append(float(r["Counter_Value"]))
    return {k: statistics.mean(v) for k, v in d.items()}, {k: len(v) for k, v in d.items()}


def hbm(a, out):
    cf, _ = counter(a.copy_fetch, lambda n: "k_copy4" in n)
    cw, _ = counter(a.copy_write, lambda n: "k_copy4" in n)
    copy_bytes = 4.0 * a.copy_floats
    f_scale = copy_bytes / (next(iter(cf.values())) * 1024)
    w_scale = copy_bytes / (next(iter(cw.values())) * 1024)
    out.write(f"Calibration: a float4 copy of {a.copy_floats} floats ({copy_bytes / 2**20:.0f} MiB read, "
              f"{copy_bytes / 2**20:.0f} MiB written) reads FETCH_SIZE {next(iter(cf.values())):.1f} KiB and "
              f"WRITE_SIZE {next(iter(cw.values())):.1f} KiB per dispatch: bytes = FETCH_SIZE KiB x 1024 x "
              f"{f_scale:.3f}, WRITE_SIZE KiB x 1024 x {w_scale:.3f}.\n\n")
    fe, n = counter(a.fetch, lambda s: re.search(r"k_(fwd|bwd)_", s) is not None)
    wr, _ = counter(a.write, lambda s: re.search(r"k_(fwd|bwd)_", s) is not None)
    out.write("| kernel | blocks | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read MB | HBM write MB |\n"
              "|---|---|---|---|---|---|---|\n")
    for k in sorted(fe):
        rd = fe[k] * 1024 * f_scale / 1e6
        wt = wr.get(k, 0.0) * 1024 * w_scale / 1e6
        out.write(f"| `{k[0]}` | {k[1]} | {n[k]} | {fe[k]:.1f} | {wr.get(k, 0.0):.1f} | {rd:.3f} | {wt:.3f} |\n")
    if a.select:
        tot = cnt = 0.0
        name = None
        for sel in a.select:
            v, b = sel.rsplit(":", 1)
            k = (v, int(b))
            byts = (fe[k] * f_scale + wr.get(k, 0.0) * w_scale) * 1024
            tot += byts * n[k]
            cnt += n[k]
            name = v.split("<")[0]
        out.write(f"\ntraffic_bytes_per_launch: {name} {tot / cnt:.0f}\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--kernels", default=r"k_(fwd|bwd)_")
    st = sub.add_parser("stats")
    st.add_argument("csv")
    st.add_argument("--top", type=int, default=16)
    h = sub.add_parser("hbm")
    for k in ("copy_fetch", "copy_write", "fetch", "write"):
        h.add_argument(k)
    h.add_argument("--copy-floats", type=int, required=True)
    h.add_argument("--select", nargs="*", default=[])
    a = ap.parse_args(argv)
    if a.mode == "trace":
        trace(a.csv, sys.stdout, a.kernels)
    elif a.mode == "stats":
        stats(a.csv, sys.stdout, a.top)
    else:
        hbm(a, sys.stdout)


if __name__ == "__main__":
    main()
