"""Summaries of rocprofv3 runs of bench.py (kernel times and HBM traffic).

    rocprofv3 --kernel-trace --stats -d DIR_STATS -o run -- python3 bench.py --no-cpu --no-psd --steps 5 --warmup 1
    rocprofv3 --pmc FETCH_SIZE -d DIR_FETCH -o run -- python3 bench.py --no-cpu --no-psd --steps 2 --warmup 1 --calibrate-copy
    rocprofv3 --pmc WRITE_SIZE -d DIR_WRITE -o run -- python3 bench.py --no-cpu --no-psd --steps 2 --warmup 1 --calibrate-copy
    python -m pmmoto_amd.profiling DIR_STATS DIR_FETCH DIR_WRITE OUT.json OUT.md [--runs TEXT]

    python -m pmmoto_amd.profiling table DIR_STATS OUT.md --runs TEXT [--note TEXT] [--pmc PMC_DIR ...]

(the second form lists every kernel of a trace with the counters of any
--pmc runs, e.g. the SQ instruction counters).

(or the same three runs of ``python3 -m pmmoto_amd.surrogate.infer --calibrate-copy``
for the surrogate's inference kernels).

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what the same
counter reads for the copy of a 1 GiB buffer that ``--calibrate-copy``
runs (1 GiB read, 1 GiB written), so the traffic is in bytes of HBM.
"""

from __future__ import annotations

import glob
import json
import os
import re
import sqlite3
import sys
from collections import defaultdict

GIB = 1 << 30

# kernel symbol -> the name bench.py reports it under
BENCH_NAMES = (
    (re.compile(r"k_pass_z<"), "edt_pass_z"),
    # k_window<tile rows, halo, input element, pass-Y output element, d2 only>: the plain
    # EDT's 16-bit passes, and the int32 ones that run when a value exceeds 65535
    (re.compile(r"k_window<\d+, \d+, unsigned short, [\w:() ]*Sq16,"), "edt_pass_y"),
    (re.compile(r"k_window<\d+, \d+, unsigned short, int,"), "edt_pass_y32"),
    (re.compile(r"k_window<\d+, \d+, [\w:() ]*Sq16, int,"), "edt_pass_x"),
    (re.compile(r"k_window<\d+, \d+, int, int,"), "edt_pass_x32"),
    (re.compile(r"k_fallback_y<[\w:() ]*Sq16>"), "edt_fallback_y"),
    (re.compile(r"k_fallback_y<int>"), "edt_fallback_y32"),
    (re.compile(r"k_fallback_x<[\w:() ]*Sq16>"), "edt_fallback_x"),
    (re.compile(r"k_fallback_x<int>"), "edt_fallback_x32"),
    (re.compile(r"k_conv_first"), None),  # surrogate inference kernels: reported under their symbol
    (re.compile(r"k_conv_igemm<"), None),
)
COPY = re.compile(r"copyBuffer")


def _short(name: str) -> str:
    """Kernel symbol without its return type, arguments and namespace
    qualifiers (also those inside template arguments)."""
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"\b[A-Za-z_]\w*::", "", name)
    name = name.split("(")[0]
    return name.replace("void ", "", 1).strip()


def _db(d: str) -> str:
    found = sorted(glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True))
    if not found:
        raise FileNotFoundError(f"no rocprofv3 results database under {d}")
    return found[0]


def kernel_times(stats_dir: str) -> dict:
    con = sqlite3.connect(_db(stats_dir))
    out = {}
    for name, n, avg in con.execute("select name, count(*), avg(duration) from kernels group by name"):
        out[_short(name)] = {"calls": n, "avg_us": avg / 1e3}
    return out


def counter(pmc_dir: str, counter_name: str, reduce=None) -> dict:
    """Per kernel symbol: the counter's mean over its dispatches (or reduce(values))."""
    con = sqlite3.connect(_db(pmc_dir))
    acc = defaultdict(list)
    for k, v in con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                            (counter_name,)):
        acc[_short(k)].append(v)
    reduce = reduce or (lambda v: sum(v) / len(v))
    return {k: reduce(v) for k, v in acc.items()}


def summarize(stats_dir: str, fetch_dir: str, write_dir: str) -> dict:
    times = kernel_times(stats_dir)
    fetch = counter(fetch_dir, "FETCH_SIZE")
    write = counter(write_dir, "WRITE_SIZE")
    # the 1 GiB copy is the largest copy dispatch (small host/device copies share its symbol)
    copy_f = [v for k, v in counter(fetch_dir, "FETCH_SIZE", max).items() if COPY.search(k)]
    copy_w = [v for k, v in counter(write_dir, "WRITE_SIZE", max).items() if COPY.search(k)]
    if not copy_f or not copy_w:
        raise RuntimeError("the 1 GiB copy kernel is missing: run bench.py with --calibrate-copy")
    f_fetch = GIB / (copy_f[0] * 1024.0)  # bytes of HBM per KiB the counter reports
    f_write = GIB / (copy_w[0] * 1024.0)
    kernels = {}
    for sym in sorted(set(times) | set(fetch) | set(write)):
        hit = next(((rx, b) for rx, b in BENCH_NAMES if rx.search(sym)), None)
        if hit is None:
            continue
        bench = hit[1] or sym
        rd = fetch.get(sym, 0.0) * 1024.0 * f_fetch
        wr = write.get(sym, 0.0) * 1024.0 * f_write
        kernels[bench] = {
            "symbol": sym,
            "avg_us": round(times.get(sym, {}).get("avg_us", float("nan")), 2),
            "calls": times.get(sym, {}).get("calls", 0),
            "hbm_read_bytes_per_launch": int(rd),
            "hbm_write_bytes_per_launch": int(wr),
            "hbm_bytes_per_launch": int(rd + wr),
        }
    return {"scale": {"fetch_bytes_per_counter_kib": f_fetch * 1024.0, "write_bytes_per_counter_kib": f_write * 1024.0},
            "kernels": kernels, "all_kernels_us": {k: round(v["avg_us"], 2) for k, v in times.items()}}


def to_markdown(s: dict, command: str) -> str:
    lines = [f"Command: `{command}`", "",
             f"Counter scale (from a 1 GiB copy): FETCH_SIZE x{s['scale']['fetch_bytes_per_counter_kib'] / 1024:.3f}, "
             f"WRITE_SIZE x{s['scale']['write_bytes_per_counter_kib'] / 1024:.3f} (bytes per reported byte)", "",
             "| kernel | avg us | HBM read GB | HBM write GB | HBM GB/s |", "|---|---|---|---|---|"]
    for b, k in s["kernels"].items():
        tot = k["hbm_bytes_per_launch"]
        rate = tot / (k["avg_us"] * 1e-6) / 1e9 if k["avg_us"] == k["avg_us"] and k["avg_us"] > 0 else float("nan")
        lines.append(f"| {b} (`{k['symbol']}`) | {k['avg_us']:.1f} | {k['hbm_read_bytes_per_launch'] / 1e9:.3f} | "
                     f"{k['hbm_write_bytes_per_launch'] / 1e9:.3f} | {rate:.0f} |")
    lines += ["", "All kernels (avg us per launch):", ""]
    lines += [f"- `{k}`: {v}" for k, v in sorted(s["all_kernels_us"].items(), key=lambda kv: -kv[1])]
    return "\n".join(lines) + "\n"


def all_counters(pmc_dir: str) -> dict:
    """kernel symbol -> {counter: mean over its dispatches} for every counter of a --pmc run."""
    con = sqlite3.connect(_db(pmc_dir))
    acc = defaultdict(lambda: defaultdict(list))
    for k, c, v in con.execute("select kernel_name, counter_name, value from counters_collection"):
        acc[_short(k)][c].append(v)
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def table(argv) -> int:
    """python -m pmmoto_amd.profiling table STATS_DIR OUT.md --runs TEXT [--note TEXT] [--match REGEX]
    [--pmc PMC_DIR ...]: a kernel table (launches, mean and total time from the kernel trace) with the
    per-dispatch mean of every counter the PMC runs recorded; FETCH_SIZE / WRITE_SIZE become HBM
    bytes when the runs hold the 1 GiB calibration copy (bench.py --calibrate-copy)."""
    import argparse
    ap = argparse.ArgumentParser(prog="pmmoto_amd.profiling table")
    ap.add_argument("stats")
    ap.add_argument("out")
    ap.add_argument("--runs", required=True, help="the rocprofv3 commands that made the directories")
    ap.add_argument("--note", default="")
    ap.add_argument("--match", default=".", help="kernels to list (regular expression on the symbol)")
    ap.add_argument("--pmc", nargs="*", default=[], help="directories of rocprofv3 --pmc runs")
    a = ap.parse_args(argv)
    times = kernel_times(a.stats)
    ctr = defaultdict(dict)
    scale = {}
    for d in a.pmc:
        for k, cs in all_counters(d).items():
            ctr[k].update(cs)
        for name in ("FETCH_SIZE", "WRITE_SIZE"):
            copies = [v for k, v in counter(d, name, max).items() if COPY.search(k)] if _has(d, name) else []
            if copies:
                scale[name] = GIB / (copies[0] * 1024.0)
    rx = re.compile(a.match)
    names = sorted((k for k in times if rx.search(k)), key=lambda k: -times[k]["avg_us"] * times[k]["calls"])
    cols = sorted({c for k in names for c in ctr.get(k, {})})
    head = ["kernel", "launches", "avg us", "total ms"] + cols
    if "FETCH_SIZE" in scale and "WRITE_SIZE" in scale:
        head += ["HBM GB/launch", "HBM GB/s"]
    shown, skip = [], 0
    for i, t in enumerate(argv):  # the --runs / --note texts are shown above and below, not twice
        if skip:
            skip -= 1
            continue
        if t in ("--runs", "--note"):
            shown.append(f"{t} '…'")
            skip = 1
        else:
            shown.append(t)
    lines = [f"Runs (from /tmp, TMPDIR=/tmp): {a.runs}", "",
             f"Command: `python -m pmmoto_amd.profiling table {' '.join(shown)}`", ""]
    if a.note:
        lines += [a.note, ""]
    if scale:
        lines += ["Counter scale (bytes per reported KiB, from the 1 GiB copy): " +
                  ", ".join(f"{k} x{v / 1024 * 1024:.1f}" for k, v in sorted(scale.items())), ""]
    lines += ["| " + " | ".join(head) + " |", "|" + "---|" * len(head)]
    for k in names:
        t = times[k]
        row = [f"`{k}`", str(t["calls"]), f"{t['avg_us']:.1f}", f"{t['avg_us'] * t['calls'] / 1e3:.2f}"]
        cs = ctr.get(k, {})
        row += [f"{cs[c]:.4g}" if c in cs else "" for c in cols]
        if "FETCH_SIZE" in scale and "WRITE_SIZE" in scale:
            if "FETCH_SIZE" in cs and "WRITE_SIZE" in cs:
                gb = (cs["FETCH_SIZE"] * 1024 * scale["FETCH_SIZE"] + cs["WRITE_SIZE"] * 1024 * scale["WRITE_SIZE"]) / 1e9
                row += [f"{gb:.3f}", f"{gb / (t['avg_us'] * 1e-6):.0f}"]
            else:
                row += ["", ""]
        lines.append("| " + " | ".join(row) + " |")
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    return 0


def _has(pmc_dir: str, name: str) -> bool:
    con = sqlite3.connect(_db(pmc_dir))
    return con.execute("select count(*) from counters_collection where counter_name = ?", (name,)).fetchone()[0] > 0


def main(argv) -> int:
    if len(argv) > 1 and argv[1] == "table":
        return table(argv[2:])
    runs = None
    if len(argv) == 8 and argv[6] == "--runs":
        runs = argv[7]
    elif len(argv) != 6:
        print(__doc__)
        return 2
    s = summarize(argv[1], argv[2], argv[3])
    with open(argv[4], "w") as f:
        json.dump(s, f, indent=1)
    md = to_markdown(s, "python -m pmmoto_amd.profiling " + " ".join(argv[1:6]) + (" --runs '…'" if runs else ""))
    with open(argv[5], "w") as f:
        f.write((f"Runs (from /tmp, TMPDIR=/tmp): {runs}\n\n" if runs else "") + md)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
