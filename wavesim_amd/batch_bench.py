"""Throughput of the two batching modes against one-problem-at-a-time solves.

    python -m wavesim_amd.batch_bench ragged [--count 32]   # independent 2-D problems, ragged grids
    python -m wavesim_amd.batch_bench rhs [--rhs 64 --n 192]  # several sources on one 3-D medium
    python -m wavesim_amd.batch_bench rhs --trace-iterations 24  # the batched solver only (counters)

Solves return once their results are final (no extra synchronisation is
needed around the clocks).  Each mode times the solves only (media, sources and plans set up before the
clock starts, every input resident on the GPU), checks that every problem
runs the same iterations as its single solve, and prints one JSON line.
The reference runs such sets as a loop of exec calls (examples/pulse2d.m:31-40).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def ragged(count: int, boundary: int, reps: int):
    from . import WaveSim, SolverGroup
    from .workloads import ragged_2d
    probs = ragged_2d(count, boundary)
    sims, svs = [], []
    for nmap, opt, src in probs:
        sim = WaveSim(nmap, opt)
        sv = sim.solver(1)
        p = sim._prepare(src)
        sv.set_source(0, p.positions, p.values)
        sims.append(sim)
        svs.append(sv)
    mi = [s.max_iterations() for s in sims]
    th = [s.energy_threshold for s in sims]
    print(f"# {count} problems, grids {sorted(set(s.N[0] for s in sims))}", file=sys.stderr, flush=True)
    for sv, m, t in zip(svs, mi, th):  # warm-up (graph capture) of the single solves
        sv.solve(m, t)
    single = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for sv, m, t in zip(svs, mi, th):
            sv.solve(m, t)
        single.append(time.perf_counter() - t0)
    its_single = [sv.result(0).iterations for sv in svs]
    print(f"# sequential {min(single):.4f} s", file=sys.stderr, flush=True)
    group = []
    with SolverGroup(svs) as g:
        g.solve(mi, th)
        for _ in range(reps):
            t0 = time.perf_counter()
            g.solve(mi, th)
            group.append(time.perf_counter() - t0)
        its_group = [sv.result(0).iterations for sv in svs]
    assert its_group == its_single, (its_group, its_single)
    upd = sum(int(np.prod(s.N)) * k for s, k in zip(sims, its_single))
    return {"mode": "ragged", "problems": count, "grids": [int(s.N[0]) for s in sims],
            "iterations": its_single, "sequential_s": round(min(single), 5), "group_s": round(min(group), 5),
            "speedup": round(min(single) / min(group), 3),
            "voxel_updates_per_s_group": round(upd / min(group), 1),
            "voxel_updates_per_s_sequential": round(upd / min(single), 1)}


def multi_rhs(nrhs: int, n: int, boundary: int, reps: int, trace_iterations: int = 0):
    from . import WaveSim
    from .simulation import Source
    from .workloads import disordered_3d
    nmap, opt, _ = disordered_3d(n, boundary)
    rng = np.random.default_rng(5)
    srcs = [Source(1.0, [int(x) for x in rng.integers(n // 4, 3 * n // 4, size=3)]) for _ in range(nrhs)]
    sim = WaveSim(nmap, opt)
    mi = sim.max_iterations()
    prepared = [sim._prepare(s) for s in srcs]
    batch = sim.solver(nrhs)
    for r, p in enumerate(prepared):
        batch.set_source(r, p.positions, p.values)
    if trace_iterations > 0:
        # only the batched solver, a fixed number of iterations (counter collection)
        batch.solve(trace_iterations, 0.0)
        return {"mode": "rhs-trace", "rhs": nrhs, "grid": [int(x) for x in sim.N],
                "iterations": batch.result(0).iterations, "plan": batch.plan}
    one = WaveSim(nmap, opt).solver(1)
    single, its_single = [], []
    for p in prepared[:1]:  # warm-up
        one.set_source(0, p.positions, p.values)
        one.solve(mi, sim.energy_threshold)
    for rep in range(reps):
        tot = 0.0
        its_single = []
        for p in prepared:
            one.set_source(0, p.positions, p.values)
            t0 = time.perf_counter()
            one.solve(mi, sim.energy_threshold)
            tot += time.perf_counter() - t0
            its_single.append(one.result(0).iterations)
        single.append(tot)
    batch.solve(mi, sim.energy_threshold)
    multi = []
    for _ in range(reps):
        t0 = time.perf_counter()
        batch.solve(mi, sim.energy_threshold)
        multi.append(time.perf_counter() - t0)
    its_multi = [batch.result(r).iterations for r in range(nrhs)]
    assert its_multi == its_single, (its_multi, its_single)
    return {"mode": "rhs", "rhs": nrhs, "grid": [int(x) for x in sim.N], "iterations": its_single,
            "sequential_s": round(min(single), 5), "batched_s": round(min(multi), 5),
            "speedup": round(min(single) / min(multi), 3)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["ragged", "rhs"])
    ap.add_argument("--count", type=int, default=32)
    ap.add_argument("--rhs", type=int, default=64)
    ap.add_argument("--n", type=int, default=192)
    ap.add_argument("--boundary", type=int, default=32)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--trace-iterations", type=int, default=0,
                    help="rhs mode: run only the batched solver for this many iterations and exit "
                         "(counter collection)")
    a = ap.parse_args(argv)
    if a.mode == "ragged":
        out = ragged(a.count, a.boundary, a.reps)
    else:
        out = multi_rhs(a.rhs, a.n, a.boundary, a.reps, a.trace_iterations)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
