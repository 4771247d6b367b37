"""Plan-specialised pass kernels (FFT plan as compile-time constants).

The shipped tuned plans get kernels whose stage radices, spans, tile width and
workgroup size are constants; every other shape runs the run-time-plan kernel.
Both compute the same butterflies in the same order, so a solve gives the same
iteration count and fields equal to rounding (the compiler may contract
multiply-adds differently in the two builds).  Bar: relative L2 <= 1e-6
between the two.  Against the oracle, the full-size tests of these workloads
(test_gpu_fullsize.py: the 320^3 fixture, 1152^2, 320^2 and the first 48
iterations at 192^3) run on the plan-specialised kernels, as the launch
counts asserted here show.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

# workload -> (pass launches of one iteration with plan-specialised kernels,
# bound on the relative L2 distance to the run-time-plan kernels' field)
# (3-D: fused first axis, x forward, z middle, x inverse; 2-D: fused, middle).
# The 320^3 and 1152^2 plans use composite register radices (320^3 x: 16,20,
# z: 20,16; 1152^2 8,12,12 / 12,12,8), which only the plan-specialised
# kernels run: without them the solver keeps its default radix order, so those
# comparisons span two factorisations and every iteration of their rounding
# differences.
SHIPPED = {"disordered_3d_256": (4, 5e-6), "spheres_3d_128": (4, 1e-6), "disordered_2d_1024": (2, 5e-6),
           "homogeneous_2d_256": (2, 1e-6)}


def _solve(name, fixed: bool, monkeypatch):
    import torch
    from wavesim_amd import WaveSim
    from wavesim_amd.workloads import WORKLOADS
    monkeypatch.setenv("WSIM_FIXED", "1" if fixed else "0")
    nmap, opt, source = WORKLOADS[name]()
    sim = WaveSim(nmap, opt)
    sv = sim.solver(1)
    src = sim._prepare(source)
    sv.set_source(0, src.positions, src.values)
    nfixed = sv.fixed_launches
    sv.solve(sim.max_iterations(), sim.energy_threshold)
    torch.cuda.synchronize()
    res = sv.result(0)
    E = sv.field(0)
    sv.close()
    return nfixed, res, E


@pytest.mark.parametrize("name", sorted(SHIPPED))
def test_fixed_kernels_match_run_time_plan_kernels(gpu, name, monkeypatch):
    nf, rf, Ef = _solve(name, True, monkeypatch)
    ng, rg, Eg = _solve(name, False, monkeypatch)
    print(f"{name}: fixed launches {nf}, iterations {rf.iterations} / {rg.iterations}, "
          f"rel-L2 {rel_l2(Ef, Eg):.2e}", flush=True)
    nexp, tol = SHIPPED[name]
    assert nf == nexp
    assert ng == 0
    assert rf.iterations == rg.iterations
    assert rel_l2(Ef, Eg) <= tol
    assert np.allclose(rf.diff_energy, rg.diff_energy, rtol=1e-5)


def test_complex128_fixed_kernels_match_run_time_plan_kernels(gpu, monkeypatch):
    """complex128 solves on a shipped plan's shape (the 256^2 homogeneous
    workload, 320^2 grid) run plan-specialised double kernels: the same
    iterations and the field of the run-time-plan double kernels to 1e-12."""
    import torch
    from wavesim_amd import WaveSim
    from wavesim_amd.workloads import WORKLOADS
    nmap, opt, source = WORKLOADS["homogeneous_2d_256"]()
    opt = dict(opt, single_precision=False)
    out = []
    for mode in ("1", "0"):
        monkeypatch.setenv("WSIM_FIXED", mode)
        sim = WaveSim(nmap, opt)
        sv = sim.solver(1)
        src = sim._prepare(source)
        sv.set_source(0, src.positions, src.values)
        nf = sv.fixed_launches
        sv.solve(sim.max_iterations(), sim.energy_threshold)
        torch.cuda.synchronize()
        out.append((nf, sv.result(0).iterations, sv.field(0)))
        sv.close()
    (nf1, it1, E1), (nf0, it0, E0) = out
    print(f"complex128 320^2: fixed launches {nf1}, iterations {it1} / {it0}, rel-L2 {rel_l2(E1, E0):.2e}", flush=True)
    assert E1.dtype == np.complex128
    assert nf1 == 2 and nf0 == 0
    assert it1 == it0
    assert rel_l2(E1, E0) <= 1e-12


def test_complex128_320_cubed_runs_its_own_plan(gpu, monkeypatch):
    """The 320^3 workload in complex128 takes the complex128 plan shipped for
    its shape (not the complex64 one), on plan-specialised double kernels; its
    field equals the run-time-plan double kernels' on the default plan (another
    factorisation, so double rounding over 360 iterations) to 1e-11, with the
    same iterations."""
    import json
    import torch
    from wavesim_amd import WaveSim, plans
    from wavesim_amd.workloads import WORKLOADS
    nmap, opt, source = WORKLOADS["disordered_3d_256"]()
    opt = dict(opt, single_precision=False)
    table = json.load(open(plans.SHIPPED))
    out = []
    for mode in ("1", "0"):
        monkeypatch.setenv("WSIM_FIXED", mode)
        sim = WaveSim(nmap, opt)
        sv = sim.solver(1)
        want = table[plans.key(sim.N, 1, len(sim.wiggles), f64=True)]["plan"]
        src = sim._prepare(source)
        sv.set_source(0, src.positions, src.values)
        out.append((sv.plan, sv.fixed_launches))
        sv.solve(sim.max_iterations(), sim.energy_threshold)
        torch.cuda.synchronize()
        out[-1] += (sv.result(0).iterations, sv.field(0))
        sv.close()
    (p1, nf1, it1, E1), (p0, nf0, it0, E0) = out
    print(f"complex128 320^3: plan {p1} ({nf1} fixed launches), {it1} / {it0} iterations "
          f"(run-time-plan kernels on {p0}), rel-L2 {rel_l2(E1, E0):.2e}", flush=True)
    assert p1 == want and want != table[plans.key(sim.N, 1, len(sim.wiggles))]["plan"]
    assert nf1 == 4 and nf0 == 0 and p0 != want
    assert E1.dtype == np.complex128 and it1 == it0
    assert rel_l2(E1, E0) <= 1e-11
