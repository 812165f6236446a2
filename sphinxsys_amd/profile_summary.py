"""Summaries of rocprofv3 CSV output restricted to a window of the run.

bench.py --profile-window prints "profile-window NAME T0 T1" (CLOCK_BOOTTIME ns,
the clock of rocprofv3's dispatch timestamps) around its headline window and
its copy calibration.  This module keeps the dispatches that start and end
inside a window and writes a Markdown summary:

* kernel times from a `--kernel-trace` run (calls, mean/min/max, share);
* per-dispatch means of the counters of any number of `--pmc` runs, and the
  figures derived from them: instructions per wave, VALU issue utilisation
  4 SQ_INSTS_VALU / (SIMDs x clock x duration) (a wave64 VALU instruction
  occupies a 16-lane SIMD for 4 cycles), L2 hit rate, HBM bytes from
  FETCH_SIZE / WRITE_SIZE scaled by the plain copies of the copy calibration;
* "traffic STAGE BYTES" lines for the acoustic passes (read by bench.py).

Dispatches of one kernel name can be told apart by their order within each
repetition (--alternate k_relation_tile=inner,contact: the first dispatch of
each advection step builds the inner lists, the second the wall contact).

Each rocprofv3 run is its own bench.py run with its own windows: every
directory is given with the log that holds its profile-window lines.

    python -m sphinxsys_amd.profile_summary --trace T=T.log --pmc SQ=SQ.log \\
        --pmc FETCH=FETCH.log --pmc WRITE=WRITE.log --out summary.md
"""
import argparse
import collections
import csv
import glob
import os
import re
import sys

SIMDS = 256 * 4          # MI355X: 256 CUs of 4 SIMDs
CLOCK_HZ = 2.4e9         # peak engine clock
COPY_BYTES = 1 << 30     # bench.py copy calibration: 1 GiB device-to-device copies
STAGES = {"acoustic_pass1": r"k_pass1<3, 1, true, true, false>", "acoustic_pass2": r"k_pass2<3, 1, true, true, false>"}


def short_name(name):
    """The kernel's name and template arguments, without namespaces and
    parameter list."""
    name = re.sub(r"sphx::\(anonymous namespace\)::", "", name)
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for n, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and n > 0:
            return name[:n].strip()
    return name.strip()


def read_windows(path):
    spans = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4 and parts[0] == "profile-window":
                spans[parts[1]] = (int(parts[2]), int(parts[3]))
    return spans


def find_csv(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit(f"profile_summary: no *{suffix} under {d}")
    return files[0]


def in_window(row, span):
    return span[0] <= int(row["Start_Timestamp"]) and int(row["End_Timestamp"]) <= span[1]


def label_dispatches(rows, alternate):
    """(dispatch id -> label): the kernel's short name, or with --alternate the
    name of its place in each repetition (dispatches in start order)."""
    by_kernel = collections.defaultdict(list)
    for r in rows:
        by_kernel[short_name(r["Kernel_Name"])].append((int(r["Start_Timestamp"]), r["Dispatch_Id"]))
    labels = {}
    for k, items in by_kernel.items():
        items.sort()
        names = next((v for pat, v in alternate.items() if k.startswith(pat)), None)
        for n, (_, did) in enumerate(items):
            labels[did] = f"{k} [{names[n % len(names)]}]" if names else k
    return labels


def busy_us(rows):
    """Time covered by at least one dispatch (the union of their intervals), us."""
    spans = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows)
    total, cur_s, cur_e = 0, None, None
    for s0, e0 in spans:
        if cur_e is None or s0 > cur_e:
            if cur_e is not None:
                total += cur_e - cur_s
            cur_s, cur_e = s0, e0
        else:
            cur_e = max(cur_e, e0)
    if cur_e is not None:
        total += cur_e - cur_s
    return total * 1e-3


def trace_stats(trace_dir, span, alternate):
    rows = [r for r in csv.DictReader(open(find_csv(trace_dir, "kernel_trace.csv"))) if in_window(r, span)]
    if not rows:
        raise SystemExit("profile_summary: no dispatch inside the window (clock mismatch?)")
    labels = label_dispatches(rows, alternate)
    dur = collections.defaultdict(list)
    for r in rows:
        dur[labels[r["Dispatch_Id"]]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    return dur, busy_us(rows)


def counter_means(pmc_dir, span, alternate):
    rows = [r for r in csv.DictReader(open(find_csv(pmc_dir, "counter_collection.csv"))) if in_window(r, span)]
    first = {}
    for r in rows:
        first.setdefault(r["Dispatch_Id"], r)
    labels = label_dispatches(list(first.values()), alternate)
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in rows:
        acc[labels[r["Dispatch_Id"]]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def copy_factors(pmc_runs):
    """KiB per dispatch that FETCH_SIZE and WRITE_SIZE read for a 1 GiB copy
    (each run's copy_calibration window), as byte factors (bytes = counter x factor)."""
    fetch = write = None
    for d, spans in pmc_runs:
        if "copy_calibration" not in spans:
            continue
        span = spans["copy_calibration"]
        for r in csv.DictReader(open(find_csv(d, "counter_collection.csv"))):
            if not in_window(r, span) or "copyBuffer" not in r["Kernel_Name"]:
                continue
            v = float(r["Counter_Value"])
            if r["Counter_Name"] == "FETCH_SIZE":
                fetch = max(fetch or 0.0, v)
            elif r["Counter_Name"] == "WRITE_SIZE":
                write = max(write or 0.0, v)
    return (COPY_BYTES / fetch if fetch else None, COPY_BYTES / write if write else None, fetch, write)


def run_arg(text):
    """DIR=LOG: an output directory and the bench.py log with its windows."""
    d, _, logf = text.partition("=")
    if not logf:
        raise SystemExit(f"profile_summary: {text}: expected DIR=LOG")
    return d, read_windows(logf)


def summarize(args):
    trace_dir, spans = run_arg(args.trace)
    span = spans[args.window]
    alternate = {}
    for a in args.alternate:
        k, v = a.split("=")
        alternate[k] = v.split(",")
    dur, busy = trace_stats(trace_dir, span, alternate)
    pmc_runs = [run_arg(p) for p in args.pmc]
    counters = collections.defaultdict(dict)
    for d, sp in pmc_runs:
        for k, cs in counter_means(d, sp[args.window], alternate).items():
            counters[k].update(cs)
    ff, fw, fetch_kib, write_kib = copy_factors(pmc_runs)
    total = sum(sum(v) for v in dur.values())
    wall_us = (span[1] - span[0]) * 1e-3
    out = []
    w = out.append
    w(f"Window `{args.window}`: {wall_us / 1e3:.1f} ms of host time, {total / 1e3:.1f} ms of kernel time "
      f"({sum(len(v) for v in dur.values())} dispatches); the GPU runs at least one kernel "
      f"{busy / 1e3:.1f} ms ({100 * busy / wall_us:.1f}% of the window)\n")
    w("| kernel | calls | avg us | min us | max us | share |")
    w("|---|---:|---:|---:|---:|---:|")
    order = sorted(dur, key=lambda k: -sum(dur[k]))
    for k in order[:args.top]:
        v = dur[k]
        w(f"| `{k}` | {len(v)} | {sum(v) / len(v):.1f} | {min(v):.1f} | {max(v):.1f} | {100 * sum(v) / total:.1f}% |")
    if counters:
        w("")
        if ff:
            w(f"Copy calibration (1 GiB device copies in the same runs): FETCH_SIZE reads {fetch_kib:,.0f} and "
              f"WRITE_SIZE {write_kib:,.0f} per dispatch for 1,048,576 KiB moved, so HBM bytes = FETCH_SIZE x "
              f"{ff:.0f} and WRITE_SIZE x {fw:.0f} ({ff / 1024:.2f} and {fw / 1024:.2f} KiB per counted KiB).\n")
        names = sorted({c for cs in counters.values() for c in cs})
        w("Counters, mean per dispatch:\n")
        w("| kernel | " + " | ".join(names) + " |")
        w("|---|" + "---:|" * len(names))
        for k in order:
            if k in counters:
                w(f"| `{k}` | " + " | ".join(f"{counters[k].get(c, float('nan')):,.0f}" for c in names) + " |")
        w("\nDerived (durations from the kernel trace):\n")
        w("| kernel | avg us | VALU/wave | SALU/wave | LDS/wave | VALU issue | L2 hit | HBM read MB | HBM write MB | HBM GB/s |")
        w("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
        traffic = {}
        for k in order:
            c = counters.get(k)
            if not c:
                continue
            us = sum(dur[k]) / len(dur[k])
            waves = c.get("SQ_WAVES")
            per = lambda n: f"{c[n] / waves:,.0f}" if waves and n in c else "-"
            valu = f"{100 * 4 * c['SQ_INSTS_VALU'] / (SIMDS * CLOCK_HZ * us * 1e-6):.0f}%" if "SQ_INSTS_VALU" in c else "-"
            hit = (f"{100 * c['TCC_HIT_sum'] / (c['TCC_HIT_sum'] + c['TCC_MISS_sum']):.0f}%"
                   if "TCC_HIT_sum" in c and c["TCC_HIT_sum"] + c["TCC_MISS_sum"] > 0 else "-")
            rd = c["FETCH_SIZE"] * ff if ff and "FETCH_SIZE" in c else None
            wr = c["WRITE_SIZE"] * fw if fw and "WRITE_SIZE" in c else None
            gbs = f"{(rd + wr) / (us * 1e-6) / 1e9:,.0f}" if rd is not None and wr is not None else "-"
            w(f"| `{k}` | {us:.1f} | {per('SQ_INSTS_VALU')} | {per('SQ_INSTS_SALU')} | {per('SQ_INSTS_LDS')} | "
              f"{valu} | {hit} | {rd / 1e6 if rd is not None else float('nan'):,.0f} | "
              f"{wr / 1e6 if wr is not None else float('nan'):,.0f} | {gbs} |")
            for stage, pat in STAGES.items():
                if k == pat and rd is not None and wr is not None:
                    traffic[stage] = rd + wr
        if traffic:
            w("\nPer-launch HBM bytes of the acoustic passes (read by bench.py):\n")
            for stage, b in traffic.items():
                w(f"    traffic {stage} {b:.0f}")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--window", default="headline")
    ap.add_argument("--trace", required=True, help="DIR=LOG: rocprofv3 --kernel-trace output and its bench.py log")
    ap.add_argument("--pmc", action="append", default=[], help="DIR=LOG: a rocprofv3 --pmc run (repeatable)")
    ap.add_argument("--alternate", action="append", default=[], help="KERNEL=name1,name2: label dispatches by order")
    ap.add_argument("--top", type=int, default=16)
    ap.add_argument("--out", default="-")
    args = ap.parse_args(argv)
    text = summarize(args)
    if args.out == "-":
        sys.stdout.write(text)
    else:
        with open(args.out, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
