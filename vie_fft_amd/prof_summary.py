"""Summarise rocprofv3 CSV output of ``bench.py`` into markdown tables.

    python -m vie_fft_amd.prof_summary trace  <dir>/bench_kernel_trace.csv
    python -m vie_fft_amd.prof_summary traffic <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv> <copy_bytes>
    python -m vie_fft_amd.prof_summary valu <counter_collection.csv> [<trace.csv>]

``trace`` lists every kernel with its launch count and mean duration.  Krylov
launches enqueued after convergence return at once (a device flag), so the
mean is also given over the launches longer than a quarter of the median —
the same rule ``vie_kernel_time`` applies to its HIP-event timings.

``traffic`` turns per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) into bytes per
launch, each scaled by the ratio of the known size of a plain device copy
to what the same counter reads for that copy.

``valu`` lists per-launch SQ_INSTS_VALU and SQ_WAVES (the median launch, so
launches that returned at once do not count) and, with a kernel trace of the
same command, the fraction of MI355X's VALU issue rate that the mean active
launch reaches: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4
cycles at 2.4 GHz = 614.4e9 wave instructions per second.
"""
from __future__ import annotations

import csv
import statistics
import sys


def _robust(v):
    med = statistics.median(v)
    kept = [x for x in v if x >= 0.25 * med]
    return sum(kept) / len(kept), len(kept)


def trace(path):
    by, info = {}, {}
    for r in csv.DictReader(open(path)):
        k = r["Kernel_Name"]
        by.setdefault(k, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        info[k] = (r["VGPR_Count"], r["LDS_Block_Size"], r["Grid_Size_X"], r["Workgroup_Size_X"])
    total = sum(sum(v) for v in by.values())
    lines = ["| kernel | launches | mean µs (all) | mean µs (active) | active | share | VGPR | LDS B | grid | WG |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        m, n = _robust(v)
        vg, lds, gx, wg = info[k]
        lines.append(f"| `{k[:120]}` | {len(v)} | {sum(v) / len(v):.1f} | {m:.1f} | {n} | "
                     f"{100 * sum(v) / total:.1f}% | {vg} | {lds} | {gx} | {wg} |")
    return "\n".join(lines)


def _counter(path, name):
    out = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == name:
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def traffic(fetch, write, copy_fetch, copy_write, copy_bytes):
    copy_bytes = float(copy_bytes)
    scale = {}
    for name, path in (("FETCH_SIZE", copy_fetch), ("WRITE_SIZE", copy_write)):
        c = _counter(path, name)
        k = max(c, key=lambda kk: len(c[kk]))  # the repeated copy kernel
        scale[name] = copy_bytes / (statistics.median(c[k]) * 1024.0)
    f = _counter(fetch, "FETCH_SIZE")
    w = _counter(write, "WRITE_SIZE")
    lines = [f"copy calibration: FETCH_SIZE x {scale['FETCH_SIZE']:.3f}, WRITE_SIZE x {scale['WRITE_SIZE']:.3f} "
             f"(plain copy of {copy_bytes:.0f} B read + written)", "",
             "| kernel | FETCH KiB | WRITE KiB | read MB (scaled) | written MB (scaled) | traffic MB |",
             "|---|---|---|---|---|---|"]
    for k in sorted(f, key=lambda kk: -sum(f[kk])):
        if k not in w:
            continue
        fm, _ = _robust(f[k])
        wm, _ = _robust(w[k])
        rb = fm * 1024 * scale["FETCH_SIZE"] / 1e6
        wb = wm * 1024 * scale["WRITE_SIZE"] / 1e6
        lines.append(f"| `{k[:120]}` | {fm:.0f} | {wm:.0f} | {rb:.1f} | {wb:.1f} | {rb + wb:.1f} |")
    return "\n".join(lines)


VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second


def valu(counters, trace_csv=None):
    ins, waves = _counter(counters, "SQ_INSTS_VALU"), _counter(counters, "SQ_WAVES")
    dur = {}
    if trace_csv:
        by = {}
        for r in csv.DictReader(open(trace_csv)):
            by.setdefault(r["Kernel_Name"], []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        dur = {k: _robust(v)[0] for k, v in by.items()}
    lines = ["| kernel | VALU instructions / launch | waves | per wave | µs (trace, active) | VALU issue fraction |",
             "|---|---|---|---|---|---|"]
    for k in sorted(ins, key=lambda kk: -statistics.median(ins[kk])):
        if k not in waves:
            continue
        i, wv = statistics.median(ins[k]), statistics.median(waves[k])
        if i < 1e6:
            continue
        us = dur.get(k)
        frac = f"{i / VALU_ISSUE_PEAK / (us * 1e-6):.3f}" if us else ""
        lines.append(f"| `{k[:120]}` | {i:.0f} | {wv:.0f} | {i / wv:.1f} | {us:.1f} | {frac} |" if us else
                     f"| `{k[:120]}` | {i:.0f} | {wv:.0f} | {i / wv:.1f} | | |")
    return "\n".join(lines)


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        print(trace(sys.argv[2]))
    elif sys.argv[1] == "traffic":
        print(traffic(*sys.argv[2:7]))
    elif sys.argv[1] == "valu":
        print(valu(*sys.argv[2:4]))
    else:
        raise SystemExit(__doc__)
