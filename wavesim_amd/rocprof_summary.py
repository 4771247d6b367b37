"""Summaries of rocprofv3 CSV output for the MBS pass kernels.

    python -m wavesim_amd.rocprof_summary stats DIR
        per-kernel dispatch count and average / median duration from the
        kernel trace (rocprofv3 --kernel-trace --stats --output-format csv -d DIR)
    python -m wavesim_amd.rocprof_summary traffic --fetch D1 --write D2 --copy-fetch D3 --copy-write D4
        HBM bytes per launch of each pass from FETCH_SIZE / WRITE_SIZE runs
        (rocprofv3 --pmc ... --output-format csv), each scaled by the reading of
        the same counter for the 1 GiB copy of `bench.py --copy-calibration`

The pass kernels are instantiations axis_pass_kernel<CONTIG, MODE, PRO, EPI,
C, NT, EPT>; they are reported under the pass names the solver's profile()
uses.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import statistics
import sys
from collections import defaultdict

COPY_BYTES = 1 << 30

_PASS = {  # (CONTIG, MODE, PRO, EPI) -> pass name
    ("true", "0", "1", "0"): "pass_fwd_pro",
    ("false", "0", "0", "0"): "pass_fwd",
    ("false", "2", "0", "0"): "pass_mid_green",
    ("false", "1", "0", "0"): "pass_inv",
    ("true", "1", "0", "1"): "pass_inv_mix",
    ("true", "1", "0", "2"): "pass_inv_mix",
    ("true", "1", "0", "3"): "pass_inv_mix_fwd",
    ("true", "2", "1", "1"): "pass_mid_pro_epi",
    ("true", "2", "1", "2"): "pass_mid_pro_epi",
    ("true", "1", "0", "4"): "pstd_inv_update",
    ("true", "1", "0", "5"): "pstd_inv_update_fwd",
    ("true", "2", "1", "4"): "pstd_step_1d",
}


def short_name(kernel: str) -> str:
    m = re.search(r"axis_pass_kernel<\s*(\w+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)", kernel)
    if m:
        # plan-specialised kernels (FixedPlan<...>) are marked "fixed"
        tag = ",fixed" if "FixedPlan" in kernel else ""
        return _PASS.get(m.groups()[:4], "axis_pass") + f"[C={m.group(5)}{tag}]"
    if "control_cycle_kernel" in kernel:
        return "control_cycle"
    if "control_kernel" in kernel:
        return "control"
    if "copyBuffer" in kernel:
        return "copyBuffer"
    if "elementwise_kernel" in kernel and "copy" in kernel.lower():
        return "torch_copy"
    return kernel.split("(")[0][-60:]


def _one(pattern: str, d: str) -> str:
    files = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not files:
        raise FileNotFoundError(f"no {pattern} under {d}")
    return files[0]


def kernel_durations(d: str):
    """{kernel: [duration ns, ...]} from *_kernel_trace.csv."""
    out = defaultdict(list)
    with open(_one("*kernel_trace.csv", d)) as f:
        for r in csv.DictReader(f):
            out[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return out


def counter_means(d: str, counter: str, skip_idle: bool = False):
    """{kernel: (mean counter value per dispatch, dispatches)} from
    *_counter_collection.csv.  skip_idle: leave out dispatches reading below 1%
    of the kernel's largest (the launches queued after a solve stopped, which
    return at once)."""
    acc = defaultdict(list)
    with open(_one("*counter_collection.csv", d)) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    if skip_idle:
        acc = {k: [x for x in v if x >= 0.01 * max(v)] for k, v in acc.items()}
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def stats(d: str) -> str:
    """Per-kernel launch statistics.  Pass launches that returned at once (the
    launches queued after a solve stopped: below 10% of the pass's median) are
    counted in `idle` and left out of the averages."""
    dur = kernel_durations(d)
    total = sum(sum(v) for v in dur.values())
    lines = [f"{'kernel':28s} {'calls':>6s} {'avg_us':>9s} {'median_us':>9s} {'min_us':>9s} {'time%':>6s} {'idle':>5s}"]
    for k, v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
        idle = 0
        if "axis_pass_kernel" in k:
            med = statistics.median(v)
            busy = [x for x in v if x >= 0.1 * med]
            idle, v = len(v) - len(busy), busy
        lines.append(f"{short_name(k):28s} {len(v):6d} {sum(v) / len(v) / 1e3:9.2f} "
                     f"{statistics.median(v) / 1e3:9.2f} {min(v) / 1e3:9.2f} {100 * sum(v) / total:6.2f} {idle:5d}")
    return "\n".join(lines)


def _copy_reading(d: str, counter: str) -> float:
    """Counter reading (KiB per dispatch) of the 1 GiB copy launches."""
    best = 0.0
    for k, (mean, n) in counter_means(d, counter).items():
        best = max(best, mean)
    return best


def traffic(fetch: str, write: str, copy_fetch: str, copy_write: str) -> str:
    cf = _copy_reading(copy_fetch, "FETCH_SIZE")
    cw = _copy_reading(copy_write, "WRITE_SIZE")
    sf, sw = COPY_BYTES / cf, COPY_BYTES / cw  # bytes per KiB reading
    lines = [f"# copy calibration: 1 GiB copy reads FETCH_SIZE {cf:.1f} KiB -> {sf:.1f} B per unit, "
             f"WRITE_SIZE {cw:.1f} KiB -> {sw:.1f} B per unit"]
    fm, wm = counter_means(fetch, "FETCH_SIZE", True), counter_means(write, "WRITE_SIZE", True)
    rows = {}
    for k in fm:
        if k in wm and "axis_pass_kernel" in k:
            rows[short_name(k).split("[")[0]] = (fm[k][0] * sf, wm[k][0] * sw, fm[k][1])
    for name, (fb, wb, n) in rows.items():
        lines.append(f"traffic {name} {fb:.0f} {wb:.0f}")
    lines.append(f"# {'pass':16s} {'fetch_MB':>10s} {'write_MB':>10s} {'dispatches':>10s}")
    for name, (fb, wb, n) in rows.items():
        lines.append(f"# {name:16s} {fb / 1e6:10.1f} {wb / 1e6:10.1f} {n:10d}")
    return "\n".join(lines)


def counters(dirs, trace: str) -> str:
    """Mean per dispatch (millions, summed over the device) of every counter
    collected in `dirs` for the per-iteration passes, the kernel-trace average
    duration from `trace`, and VALU busy = SQ_INSTS_VALU x 4 cycles (a wave64
    VALU instruction on a 16-lane SIMD) / 1024 SIMDs over that duration at
    2.4 GHz (an upper clock, so a lower bound on busy)."""
    passes = ["pass_fwd", "pass_mid_green", "pass_inv", "pass_inv_mix_fwd"]
    table = {}
    for d in dirs:
        acc = defaultdict(lambda: defaultdict(list))
        with open(_one("*counter_collection.csv", d)) as f:
            for r in csv.DictReader(f):
                name = short_name(r["Kernel_Name"]).split("[")[0]
                if name in passes:
                    acc[r["Counter_Name"]][name].append(float(r["Counter_Value"]))
        for c, per in acc.items():
            # launches queued after a stop read next to nothing
            table[c] = {n: sum(x for x in v if x >= 0.01 * max(v)) / max(1, sum(1 for x in v if x >= 0.01 * max(v)))
                        for n, v in per.items()}
    dur = {}
    for k, v in kernel_durations(trace).items():
        name = short_name(k).split("[")[0]
        if name in passes:
            med = statistics.median(v)
            busy = [x for x in v if x >= 0.1 * med]
            dur[name] = sum(busy) / len(busy) / 1e3
    lines = [f"{'counter':34s}" + "".join(f"{p:>18s}" for p in passes)]
    for c in sorted(table):
        lines.append(f"{c:34s}" + "".join(f"{table[c].get(p, float('nan')) / 1e6:18.3f}" for p in passes))
    lines.append(f"{'duration_us (kernel trace)':34s}" + "".join(f"{dur.get(p, float('nan')):18.2f}" for p in passes))
    if "SQ_INSTS_VALU" in table:
        busy = [table["SQ_INSTS_VALU"].get(p, 0) * 4 / 1024 / (dur.get(p, float("inf")) * 2400) for p in passes]
        lines.append(f"{'valu_busy':34s}" + "".join(f"{b:18.2f}" for b in busy))
    return "\n".join(lines)


def fno_forward(d: str, layers: int = 4) -> str:
    """Per-role GEMM times of the last network forward in a kernel trace (a
    forward starts with its complex_weights launch; its GEMMs follow in the
    order of wavesim_amd.learned.fno.forward_gemm_roles)."""
    from .learned.fno import Config, forward_gemm_roles
    with open(_one("*kernel_trace.csv", d)) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(rows) if "complex_weights_kernel" in r["Kernel_Name"]]
    if not starts:
        raise ValueError(f"no network forward in the trace under {d}")
    roles = forward_gemm_roles(Config(layers=layers))
    fwd = rows[starts[-1]:]
    gemms = [r for r in fwd if "gemm_kernel" in r["Kernel_Name"]][:len(roles)]
    if len(gemms) != len(roles):
        raise ValueError(f"{len(gemms)} GEMM launches after the last forward's start, {len(roles)} expected")
    dur = lambda r: (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    other = defaultdict(list)
    for r in fwd:
        if "gemm_kernel" not in r["Kernel_Name"]:
            other[r["Kernel_Name"].split("(")[0].split("::")[-1]].append(dur(r))
    acc = defaultdict(list)
    for role, r in zip(roles, gemms):
        acc[role].append(dur(r))
    total = sum(dur(r) for r in gemms)
    lines = [f"# last forward: {len(gemms)} GEMM launches {total:.1f} us, other kernels "
             + ", ".join(f"{k} {len(v)}x {sum(v):.1f} us" for k, v in other.items()),
             f"{'role':28s} {'launches':>8s} {'avg_us':>8s} {'total_us':>9s}"]
    for role, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{role:28s} {len(v):8d} {sum(v) / len(v):8.2f} {sum(v):9.1f}")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    for name in ("fetch", "write", "copy-fetch", "copy-write"):
        t.add_argument("--" + name, required=True)
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--trace", required=True)
    n = sub.add_parser("fno")
    n.add_argument("dir")
    n.add_argument("--layers", type=int, default=4)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats(a.dir))
    elif a.cmd == "counters":
        print(counters(a.dirs, a.trace))
    elif a.cmd == "fno":
        print(fno_forward(a.dir, a.layers))
    else:
        print(traffic(a.fetch, a.write, a.copy_fetch, a.copy_write))
    return 0


if __name__ == "__main__":
    sys.exit(main())
