"""Autotune the FFT plans of the benchmark workloads and print them as JSON.

    python -m wavesim_amd.autotune [workload ...] [--reps 4]
    python -m wavesim_amd.autotune --ragged     # the grids of the ragged 2-D batch

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
    ap.add_argument("--ragged", action="store_true",
                    help="tune one problem of each grid size of the ragged 2-D batch instead")
    a = ap.parse_args(argv)
    if a.ragged:
        from .workloads import ragged_2d
        sims = {}
        for nmap, opt, _ in ragged_2d():
            sim = WaveSim(nmap, opt)
            sims.setdefault(plans.key(sim.N, a.nrhs, len(sim.wiggles)), (f"ragged {sim.N[0]}x{sim.N[1]}", sim))
        todo = list(sims.values())
    else:
        todo = []
        for name in a.workloads:
            nmap, opt, _ = WORKLOADS[name]()
            todo.append((name, (WaveSimVector if name.startswith("sphere_vector") else WaveSim)(nmap, opt)))
    out = {}
    for name, sim in todo:
        sv = sim.solver(a.nrhs)
        before = sv.plan
        plan, ms = sim.tune(a.nrhs, a.reps)
        out[plans.key(sim.N, a.nrhs, len(sim.wiggles))] = {"plan": plan, "ms_per_iteration": round(ms, 4)}
        print(f"# {name}: default {before} -> {plan} ({ms:.4f} ms/iteration)", file=sys.stderr, flush=True)
        sv.close()
        sim._solver = None
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
