"""Autotune the FFT plans of the benchmark workloads and print them as JSON.

    python -m wavesim_amd.autotune [workload ...] [--reps 4]

Each plan is also stored in the user plan cache; the printed mapping is the
format of ``tuned_plans.json``.
"""
from __future__ import annotations

import argparse
import json
import sys

from . import WaveSim, WaveSimVector, plans
from .workloads import WORKLOADS


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--reps", type=int, default=4)
    ap.add_argument("--nrhs", type=int, default=1)
    a = ap.parse_args(argv)
    out = {}
    for name in a.workloads:
        nmap, opt, _ = WORKLOADS[name]()
        sim = (WaveSimVector if name.startswith("sphere_vector") else WaveSim)(nmap, opt)
        sv = sim.solver(a.nrhs)
        before = sv.plan
        import ctypes  # noqa: F401  (keeps the library loaded)
        t0 = sv._lib  # noqa: F841
        plan, ms = sim.tune(a.nrhs, a.reps)
        out[plans.key(sim.N, a.nrhs, len(sim.wiggles))] = {"plan": plan, "ms_per_iteration": round(ms, 4)}
        print(f"# {name}: default {before} -> {plan} ({ms:.4f} ms/iteration)", file=sys.stderr, flush=True)
        sv.close()
        sim._solver = None
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
