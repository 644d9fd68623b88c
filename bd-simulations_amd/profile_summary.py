"""Summaries of rocprofv3 CSV output for the committed profiles.

python bd-simulations_amd/profile_summary.py kernels DIR
    per-kernel dispatch statistics from DIR's *kernel_trace.csv: calls,
    average/min/max duration, share of GPU time, and the average over the
    dispatches that overlapped no other dispatch ("alone": the figure that
    a kernel's back-to-back HIP-event timing measures), plus the period of
    the graph-replayed steps (consecutive starts of the step's first kernel
    with every kernel of a step in between).
python bd-simulations_amd/profile_summary.py traffic FETCH_DIR WRITE_DIR [--json OUT]
    HBM bytes per dispatch from two --pmc runs (FETCH_SIZE, WRITE_SIZE, KiB
    per dispatch), each scaled by the reading of the runtime's copy kernel
    (the last --copies copy dispatches, which bench.py --kernel-probe makes
    after every other launch: 256 MiB each, every byte read and written once).
python bd-simulations_amd/profile_summary.py counters DIR
    per-kernel averages of every counter of a --pmc run (e.g. the SQ
    instruction mix).

Kernel names are shortened to the identifier after the namespace (k_wca_f32,
k_update, ...); the runtime's copy kernel keeps its own name.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short(name):
    name = name.strip()
    m = re.search(r"bdkx::(\w+)", name)
    if m:
        return m.group(1)
    m = re.match(r"\s*(?:void\s+)?([A-Za-z_][\w]*)", name)
    return m.group(1) if m else name


def rows(directory, suffix):
    files = sorted(glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit("no *%s under %s" % (suffix, directory))
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                yield r


STEP_KERNELS = ("k_update", "k_rank_scan", "k_fill", "k_sort_cells", "k_wca_f32", "k_deposit", "k_r2c_rows",
                "k_col", "k_c2r_rows")


def kernels(directory, step_kernels=STEP_KERNELS):
    ev = []
    for r in rows(directory, "kernel_trace.csv"):
        ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"])))
    ev.sort()
    # a dispatch is "alone" when no other dispatch overlaps it in time
    alone = [True] * len(ev)
    run_end, run_idx = -1, -1
    for i, (s, e, _) in enumerate(ev):
        if s < run_end:
            alone[i] = False
            if run_idx >= 0:
                alone[run_idx] = False
        if e > run_end:
            run_end, run_idx = e, i
    total = sum(e - s for s, e, _ in ev)
    per = defaultdict(list)
    per_alone = defaultdict(list)
    for (s, e, n), a in zip(ev, alone):
        per[n].append(e - s)
        if a:
            per_alone[n].append(e - s)
    out = ["| kernel | calls | avg us | min us | max us | alone avg us (n) | % of GPU time |",
           "|---|---|---|---|---|---|---|"]
    for n, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        al = per_alone.get(n, [])
        out.append("| %s | %d | %.2f | %.2f | %.2f | %s | %.2f |" % (
            n, len(d), sum(d) / len(d) / 1e3, min(d) / 1e3, max(d) / 1e3,
            ("%.2f (%d)" % (sum(al) / len(al) / 1e3, len(al))) if al else "-", 100.0 * sum(d) / total))
    # step periods: the time between consecutive starts of the step's first
    # kernel when every kernel of a step started in between (graph-replayed
    # steps; the one-kernel timing runs between them do not qualify)
    need = set(step_kernels)
    starts = [i for i, (_, _, n) in enumerate(ev) if n == step_kernels[0]]
    periods = []
    for a, b in zip(starts, starts[1:]):
        if need <= {ev[i][2] for i in range(a, b)}:
            periods.append(ev[b][0] - ev[a][0])
    if periods:
        periods.sort()
        out.append("")
        out.append("Graph-replayed steps (consecutive %s starts enclosing all of %s): %d, median period %.2f us, "
                   "min %.2f us." % (step_kernels[0], ", ".join(step_kernels[1:]), len(periods),
                                      periods[len(periods) // 2] / 1e3, periods[0] / 1e3))
    return "\n".join(out)


def per_kernel_counter(directory):
    acc = defaultdict(lambda: defaultdict(list))
    for r in rows(directory, "counter_collection.csv"):
        acc[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def probe_copy(directory, counter, copies):
    """The counter's average over the last `copies` dispatches of the copy
    kernel: the calibration copies that bench.py --kernel-probe makes after
    every other launch (earlier copy dispatches are the program's own)."""
    ev = [(int(r["Start_Timestamp"]), float(r["Counter_Value"]))
          for r in rows(directory, "counter_collection.csv")
          if short(r["Kernel_Name"]) == COPY_KERNEL and r["Counter_Name"] == counter]
    ev.sort()
    last = [v for _, v in ev[-copies:]]
    return sum(last) / len(last) if last else None


def traffic(fetch_dir, write_dir, json_out=None, copies=3):
    f = per_kernel_counter(fetch_dir)
    w = per_kernel_counter(write_dir)
    probe_kib = 256 * 1024  # bench.py --kernel-probe copies 256 MiB each way
    cf = probe_copy(fetch_dir, "FETCH_SIZE", copies)
    cw = probe_copy(write_dir, "WRITE_SIZE", copies)
    fs = probe_kib / cf if cf else 1.0
    ws = probe_kib / cw if cw else 1.0
    out = ["Calibration: the copy kernel reads %s KiB and writes %s KiB per launch for %d KiB moved each way, "
           "so FETCH_SIZE is scaled by %.4f and WRITE_SIZE by %.4f." % (
               "%.0f" % cf if cf else "?", "%.0f" % cw if cw else "?", probe_kib, fs, ws), "",
           "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled traffic MB |", "|---|---|---|---|"]
    js = {"fetch_scale": fs, "write_scale": ws, "kernels": {}}
    for k in sorted(set(f) & set(w)):
        if k == COPY_KERNEL:
            continue
        fk, wk = f[k].get("FETCH_SIZE", 0.0), w[k].get("WRITE_SIZE", 0.0)
        mb = 1024.0 * (fk * fs + wk * ws) / 1e6
        out.append("| %s | %.0f | %.0f | %.1f |" % (k, fk, wk, mb))
        js["kernels"][k] = {"fetch_kib": fk, "write_kib": wk}
    if json_out:
        with open(json_out, "w") as fh:
            json.dump(js, fh, indent=1, sort_keys=True)
    return "\n".join(out)


def counters(directory):
    per = per_kernel_counter(directory)
    names = sorted({c for cs in per.values() for c in cs})
    out = ["| kernel | " + " | ".join(names) + " |", "|---" * (len(names) + 1) + "|"]
    for k, cs in sorted(per.items()):
        out.append("| %s | " % k + " | ".join("%.1f" % cs.get(c, float("nan")) for c in names) + " |")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("kernels")
    a.add_argument("dir")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--json")
    b.add_argument("--copies", type=int, default=3, help="calibration copies at the end of the run")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    args = ap.parse_args()
    if args.cmd == "kernels":
        print(kernels(args.dir))
    elif args.cmd == "traffic":
        print(traffic(args.fetch_dir, args.write_dir, args.json, args.copies))
    else:
        print(counters(args.dir))


if __name__ == "__main__":
    sys.exit(main())
