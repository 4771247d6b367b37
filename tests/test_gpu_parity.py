"""HIP path vs the CPU oracle (complex64 restatement of the reference loop).

Bar: relative L2 <= 1e-5 in complex64, with the same iteration
count; the every-iteration ROI energies agree to 1e-4 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests import cases
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _pair(n, opt, vector=False):
    from oracle import mbs
    import wavesim_amd as w
    o = mbs.make_sim(n, opt, vector=vector)
    g = (w.WaveSimVector if vector else w.WaveSim)(n, opt)
    return o, g


def _run_both(n, opt, src, vector=False, max_iterations=None):
    from oracle import mbs
    o, g = _pair(n, opt, vector)
    Eo, so = mbs.run(o, cases.oracle_source(src), max_iterations=max_iterations, record_every=True)
    Eg, sg = g.exec(cases.gpu_source(src), max_iterations=max_iterations)
    return o, g, Eo, so, Eg, sg


def _check_solve(Eo, so, Eg, sg, tol=TOL):
    print(f"iterations oracle={so.it - 1} gpu={sg.iterations}", flush=True)
    assert sg.iterations == so.it - 1
    assert sg.converged == so.converged
    e = rel_l2(Eg, Eo)
    print(f"rel-L2 E = {e:.3e}", flush=True)
    assert e <= tol
    ev = np.asarray(so.energy_every)
    assert np.allclose(sg.energy_every, ev, rtol=1e-4, atol=1e-4 * ev.max())
    assert np.allclose(sg.diff_energy, so.diff_energy, rtol=1e-4)


@pytest.mark.parametrize("builder", [cases.case_2d_homogeneous, cases.case_2d_disordered,
                                     cases.case_3d_disordered, cases.case_1d, cases.case_periodic_2d])
def test_propagate_scalar(gpu, builder):
    from oracle import mbs
    n, opt, _ = builder()
    o, g = _pair(n, opt)
    rng = np.random.default_rng(0)
    N = o.N
    for w in range(len(o.wiggles)):
        E = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
        ref = mbs.propagate(o, E, o.wiggles[w])
        out = g.propagate(E, w)
        print(builder.__name__, "wiggle", w, rel_l2(out, ref), flush=True)
        assert rel_l2(out, ref) <= TOL


def test_propagate_vector(gpu):
    from oracle import mbs
    n, opt, _ = cases.case_3d_vector_sphere()
    o, g = _pair(n, opt, vector=True)
    rng = np.random.default_rng(1)
    for w in range(len(o.wiggles)):
        E = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)).astype(np.complex64)
        ref = mbs.propagate(o, E, o.wiggles[w])
        out = g.propagate(E, w)
        assert rel_l2(out, ref) <= TOL


@pytest.mark.parametrize("builder", [cases.case_2d_homogeneous, cases.case_2d_disordered,
                                     cases.case_1d, cases.case_periodic_2d, cases.case_3d_disordered])
def test_solve_scalar(gpu, builder):
    n, opt, src = builder()
    _, _, Eo, so, Eg, sg = _run_both(n, opt, src)
    _check_solve(Eo, so, Eg, sg)


def _absorbing_3d_periodic_z():
    n = cases.smooth_disorder((18, 20, 16), 5, absorption=0.02)
    return n, cases.base_opt([4, 6, 0], 1e-6, **{"lambda": 0.8}), [(1.0, [9, 10, 8])], False


def _vector_2d_y_polarised():
    n = cases.smooth_disorder((24, 28), 6)
    return n, cases.base_opt([6, 6], 1e-6), [(np.ones((24, 1, 1, 1)), [1, 1, 1, 2])], True


def _vector_3d_z_polarised_point():
    n = cases.sphere((16, 16, 16), 4, 1.5, 1.2)
    return n, cases.base_opt([4, 4, 4], 1e-5), [(1.0, [8, 8, 8, 3])], True


@pytest.mark.parametrize("builder", [_absorbing_3d_periodic_z, _vector_2d_y_polarised, _vector_3d_z_polarised_point])
def test_solve_more_configurations(gpu, builder):
    """Against the oracle: a 3-D absorbing medium with a periodic z axis
    (Nw = 4 in 3-D) and wavelength 0.8, a 2-D vector field with a
    y-polarised source, a 3-D vector field with a z-polarised point source."""
    n, opt, src, vector = builder()
    o, g, Eo, so, Eg, sg = _run_both(n, opt, src, vector=vector)
    _check_solve(Eo, so, Eg, sg)


def test_solve_vector(gpu):
    n, opt, src = cases.case_3d_vector_sphere()
    _, _, Eo, so, Eg, sg = _run_both(n, opt, src, vector=True)
    _check_solve(Eo, so, Eg, sg)


def test_primed_field_full_grid(gpu):
    """The MEX output E' (full grid, un-divided) matches everywhere, including
    the boundary layers where gamma = 0."""
    n, opt, src = cases.case_2d_disordered()
    _, _, _, so, _, sg = _run_both(n, opt, src)
    assert rel_l2(sg.E_primed, so.E_primed) <= TOL


def test_run_algorithm_dropin(gpu):
    """wsim_run_algorithm: the RunWaveSimAlgorithmMex-shaped entry point."""
    import ctypes as C
    from oracle import mbs
    from wavesim_amd import capi
    n, opt, src = cases.case_2d_homogeneous()
    o = mbs.make_sim(n, opt)
    Eo, so = mbs.run(o, cases.oracle_source(src))
    s = mbs.prepare_source(o, cases.oracle_source(src))
    L = capi.lib()
    W = capi.Wiggles(o.wiggles)
    P = capi.Patches(s.positions, s.values)
    N = (C.c_double * 4)(*o.N)
    roi = (C.c_double * 8)(*np.asarray(o.roi, np.float64).T.ravel())
    g = capi.c64(o.gamma)
    Eout = np.zeros(o.N, np.complex64, order="F")
    de = (C.c_double * 100000)()
    nde = C.c_int64()
    rc = L.wsim_run_algorithm(N, capi.ptr(g), W.n, W.arr, P.n, P.arr, float(o.epsilon),
                              capi.cfloat(o.k02e.real, o.k02e.imag), roi, float("inf"),
                              o.energy_threshold, b"NoCallback", 50.0, capi.NULL_CALLBACK, None,
                              capi.ptr(Eout), de, 100000, C.byref(nde))
    capi.check(rc)
    assert nde.value == so.it - 1
    assert rel_l2(Eout, so.E_primed) <= TOL
    assert np.allclose(np.asarray(de[:nde.value]), so.diff_energy, rtol=1e-4)


def test_batched_rhs_equal_single(gpu):
    """Several sources in one batched launch give the single-source results."""
    import wavesim_amd as w
    n, opt, _ = cases.case_2d_disordered(thr=1e-5)
    sim = w.WaveSim(n, opt)
    srcs = [w.Source(1.0, [10, 12]), w.Source(np.ones(40), [1, 1]), w.Source(2.0 - 1j, [30, 40])]
    Eb, sb = sim.exec_batch(srcs)
    for k, s in enumerate(srcs):
        E1, s1 = w.WaveSim(n, opt).exec(s)
        assert sb[k].iterations == s1.iterations
        assert rel_l2(Eb[k], E1) <= 1e-6


def test_superposition(gpu):
    """Linearity in S at a fixed iteration count (pins Source.plus)."""
    import wavesim_amd as w
    n, opt, _ = cases.case_2d_disordered()
    sim = w.WaveSim(n, opt)
    a, b = w.Source(1.0, [10, 12]), w.Source(0.5j, [25, 30])
    Ea, _ = sim.exec(a, max_iterations=40)
    Eb, _ = sim.exec(b, max_iterations=40)
    Eab, _ = sim.exec(a + b, max_iterations=40)
    # the stop rule only acts at multiples of Nw; force the same count
    assert rel_l2(Eab, Ea + Eb) <= 1e-5


def test_max_iterations_cap(gpu):
    import wavesim_amd as w
    n, opt, src = cases.case_2d_homogeneous(thr=1e-30)
    E, st = w.WaveSim(n, opt).exec(cases.gpu_source(src), max_iterations=20)
    assert st.iterations == 20 and not st.converged


@pytest.mark.parametrize("interval", [1, 7, 50, 300])
def test_callback_cadence_follows_simulation_next(gpu, interval):
    """Simulation.next calls the callback after every iteration that is a
    multiple of callback_interval and after the last one (Simulation.m:213):
    ceil(iterations / interval) calls, each with the history up to its
    iteration, and the solve itself is unchanged by the callback."""
    import math
    import wavesim_amd as w
    n, opt, src = cases.case_2d_disordered()
    # callback_interval also caps the energy cadence (Simulation.m:135): the
    # solve without a callback uses the same options
    opt2 = dict(opt, callback_interval=interval)
    E0, s0 = w.WaveSim(n, opt2).exec(cases.gpu_source(src))
    calls = []
    g2 = w.WaveSim(n, opt2)
    g2.callback = lambda it, hist: calls.append((it, len(hist))) and False
    E1, s1 = g2.exec(cases.gpu_source(src))
    it = s1.iterations
    print(f"interval {interval}: {it} iterations, {len(calls)} calls", flush=True)
    assert it == s0.iterations and np.array_equal(E1, E0)
    want = [k for k in range(interval, it + 1, interval)]
    if not want or want[-1] != it:
        want.append(it)
    assert len(calls) == math.ceil(it / interval)
    assert [c[0] for c in calls] == want and all(c[0] == c[1] for c in calls)


@pytest.mark.parametrize("max_iterations", [6, 7])
def test_last_increment_on_odd_iteration(gpu, max_iterations):
    """E' takes its increments two at a time, on even iterations; a solve that
    stops after an odd iteration adds its last increment at the end.  With
    Nw = 1 (periodic, no boundary layers) a stop can fall on any iteration;
    single solves, multi-RHS solves and groups all finish E'."""
    import wavesim_amd as w
    n, opt, src = cases.case_periodic_2d(thr=1e-30)
    o, g, Eo, so, Eg, sg = _run_both(n, opt, src, max_iterations=max_iterations)
    assert sg.iterations == max_iterations
    _check_solve(Eo, so, Eg, sg)
    [(Em, sm)] = w.exec_many([(w.WaveSim(n, opt), cases.gpu_source(src))], max_iterations=max_iterations)
    assert sm.iterations == max_iterations and np.array_equal(Em, Eg)
    Eb, sb = g.exec_batch([cases.gpu_source(src)] * 2, max_iterations=max_iterations)
    assert all(np.array_equal(e, Eg) for e in Eb)


def test_iteration_cap_waits_for_a_full_wiggle_cycle(gpu):
    """max_iterations takes effect at the first multiple of Nw at or after it
    (Simulation.m:202): a fresh solver capped at 18 with Nw = 4 runs 20
    iterations, as the oracle does."""
    n, opt, src = cases.case_2d_homogeneous(thr=1e-30)
    o, g, Eo, so, Eg, sg = _run_both(n, opt, src, max_iterations=18)
    assert so.it - 1 == 20
    _check_solve(Eo, so, Eg, sg)


@pytest.mark.parametrize("builder,vector", [(cases.case_2d_disordered, False), (cases.case_3d_disordered, True)])
def test_problem_built_by_the_library_solves_like_exec(gpu, builder, vector):
    """A solver created from wsim_problem_build's outputs (the C++ set-up, no
    Python mirror) gives exec's field and energies bit for bit."""
    import wavesim_amd as w
    from wavesim_amd import problem
    from wavesim_amd.solver import MBSSolver
    n, opt, src = builder()
    sim = (w.WaveSimVector if vector else w.WaveSim)(n, opt)
    Eg, sg = sim.exec(cases.gpu_source(src))
    p = problem.build(n, opt, vector=vector)
    sv = MBSSolver(p["N"], p["gamma"], p["wiggles"], float(p["epsilon"]), complex(p["k02e"]), p["roi"],
                   sim.energy_calculation_interval)
    ps = sim._prepare(cases.gpu_source(src))
    sv.set_source(0, ps.positions, ps.values)
    sv.solve(sim.max_iterations(), sim.energy_threshold)
    r = sv.result(0)
    assert r.iterations == sg.iterations
    assert np.array_equal(r.diff_energy, sg.diff_energy)
    with np.errstate(invalid="ignore", divide="ignore"):  # gamma is 0 at the grid's outer edge
        E = sim.crop_field(sv.field(0) / np.asarray(sim.gamma)[..., None])
    assert np.array_equal(E, Eg)
    sv.close()


def test_zero_source(gpu):
    import wavesim_amd as w
    n, opt, _ = cases.case_2d_homogeneous()
    E, st = w.WaveSim(n, opt).exec(w.Source(0.0, [3, 3]))
    assert st.iterations == 0 and np.all(E == 0)


def test_result_before_any_solve(gpu):
    """A fresh handle reports no iterations, not converged and an empty history."""
    import wavesim_amd as w
    n, opt, _ = cases.case_2d_homogeneous()
    sv = w.WaveSim(n, opt).solver(3)
    for r in range(3):
        res = sv.result(r)
        assert res.iterations == 0 and not res.converged and res.diff_energy.size == 0


def test_deterministic(gpu):
    import wavesim_amd as w
    n, opt, src = cases.case_2d_disordered()
    sim = w.WaveSim(n, opt)
    E1, s1 = sim.exec(cases.gpu_source(src))
    E2, s2 = sim.exec(cases.gpu_source(src))
    assert np.array_equal(E1, E2) and np.array_equal(s1.diff_energy, s2.diff_energy)


@pytest.mark.parametrize("shape", [(42, 30), (44, 22), (54, 21), (25, 49)])
def test_odd_radices(gpu, shape):
    """Axis lengths with factors 3, 5, 7 and 11 (periodic so N is the map size)."""
    from oracle import mbs
    nmap = cases.smooth_disorder(shape, 7, absorption=0.05)
    opt = cases.base_opt([0, 0], 1e-8)
    o, g = _pair(nmap, opt)
    rng = np.random.default_rng(2)
    E = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)).astype(np.complex64)
    assert rel_l2(g.propagate(E, 0), mbs.propagate(o, E, o.wiggles[0])) <= TOL


@pytest.mark.parametrize("shape", [(16, 2112), (2112, 8), (24, 3072)])
def test_long_axes(gpu, shape):
    """Axes whose one-column tile with its tables exceeds the usual per-workgroup
    LDS target (2112 = 2^6 3 11 strided and contiguous; 3072 strided)."""
    from oracle import mbs
    nmap = cases.smooth_disorder(shape, 4, absorption=0.05)
    opt = cases.base_opt([0, 0], 1e-8)
    o, g = _pair(nmap, opt)
    rng = np.random.default_rng(3)
    E = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)).astype(np.complex64)
    for w in range(len(o.wiggles)):
        assert rel_l2(g.propagate(E, w), mbs.propagate(o, E, o.wiggles[w])) <= TOL


@pytest.mark.parametrize("builder", [cases.case_2d_disordered, cases.case_3d_disordered])
def test_fused_first_pass_matches_separate_passes(gpu, builder, monkeypatch):
    """The last inverse pass running the next iteration's first forward pass
    (the default) computes the same operations as separate passes; only the
    compiler's multiply-add contraction may differ (last bits)."""
    import wavesim_amd as w
    n, opt, src = builder()
    E1, s1 = w.WaveSim(n, opt).exec(cases.gpu_source(src))
    monkeypatch.setenv("WSIM_FUSE", "0")
    E0, s0 = w.WaveSim(n, opt).exec(cases.gpu_source(src))
    assert s1.iterations == s0.iterations
    assert rel_l2(E1, E0) <= 1e-6
    assert np.allclose(s1.diff_energy, s0.diff_energy, rtol=1e-5)


@pytest.mark.parametrize("builder", [cases.case_2d_disordered, cases.case_3d_disordered])
def test_propagate_matches_rocfft(gpu, builder):
    """The hand-written FFT passes against rocFFT: the same wiggle ramps and
    scalar Green's function (tables from the oracle) around torch.fft.fftn /
    ifftn on the GPU, in complex64."""
    import torch
    n, opt, src = builder()
    o, g = _pair(n, opt)
    rng = np.random.default_rng(8)
    E = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)).astype(np.complex64)
    for w, wig in enumerate(o.wiggles):
        t = lambda a, shape: torch.from_numpy(np.asarray(a).astype(np.complex64).reshape(shape)).cuda()
        ramp = t(wig["gy"], (-1, 1, 1, 1)) * t(wig["gx"], (1, -1, 1, 1)) * t(wig["gz"], (1, 1, -1, 1))
        den = ((wig["pxe"] ** 2).reshape(1, -1, 1, 1) + (wig["pye"] ** 2).reshape(-1, 1, 1, 1)
               + (wig["pze"] ** 2 - o.k02e).reshape(1, 1, -1, 1))
        x = torch.from_numpy(E).cuda() * ramp
        x = torch.fft.ifftn(torch.fft.fftn(x, dim=(0, 1, 2)) / t(den, den.shape), dim=(0, 1, 2))
        ref = (x * ramp.conj()).cpu().numpy()
        assert rel_l2(g.propagate(E, w), ref) <= TOL


def test_1d_against_reference_analytic_solution(gpu):
    """The GPU solve of a 1-D homogeneous line with ARL + ACC against the
    reference's utilities/homogeneous_medium_analytic_solution.m (restated in
    oracle.analytic), in complex64 to an energy ratio of 1e-10."""
    import wavesim_amd as w
    from oracle import analytic
    Ny = 256
    sim = w.WaveSim(np.ones((Ny, 1)), dict(pixel_size=0.25, boundary_widths=[32, 0], energy_threshold=1e-10))
    E, st = sim.exec(w.Source(1.0, [Ny // 2, 1]))
    x = (np.arange(Ny) - (Ny // 2 - 1)) * 0.25
    ref = analytic.homogeneous_1d(sim.k0, 0.25, x)
    e = rel_l2(np.asarray(E).reshape(-1), ref)
    print(f"1-D analytic rel-L2 {e:.2e} after {st.iterations} iterations", flush=True)
    assert st.converged and e < 2e-5  # the complex64 oracle reaches 7.0e-6 here


def test_custom_roi(gpu):
    n, opt, src = cases.case_2d_disordered()
    opt = dict(opt, roi=[[5, 7], [30, 44]])
    _, _, Eo, so, Eg, sg = _run_both(n, opt, src)
    _check_solve(Eo, so, Eg, sg)


def test_plan_roundtrip_and_tuned_parity(gpu):
    """A tuned FFT plan (radix order, tile width) changes speed, not results."""
    from oracle import mbs
    n, opt, src = cases.case_3d_disordered()
    o, g = _pair(n, opt)
    sv = g.solver()
    default = sv.plan
    plan, ms = sv.tune(2)
    assert ms > 0 and plan
    sv.plan = "y:4,7/0;x:16,2/8;z:8,3/16"
    assert sv.plan == "y:4,7/0;x:16,2/8;z:8,3/16"
    rng = np.random.default_rng(5)
    E = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)).astype(np.complex64)
    for p in (default, plan, "y:4,7/0;x:16,2/8;z:8,3/16", "y:2,2,7/4;x:2,4,4/32;z:3,8/8"):
        sv.plan = p
        for w in (0, 5):
            assert rel_l2(g.propagate(E, w), mbs.propagate(o, E, o.wiggles[w])) <= TOL, p
    Eo, so = mbs.run(o, cases.oracle_source(src))
    sv.plan = plan
    Eg, sg = g.exec(cases.gpu_source(src))
    assert sg.iterations == so.it - 1 and rel_l2(Eg, Eo) <= TOL


@pytest.mark.parametrize("builder", [cases.case_2d_disordered, cases.case_3d_disordered])
def test_xcd_tile_ranges_change_nothing(gpu, builder):
    """Handing each XCD one contiguous range of tiles permutes the workgroups
    only: fields and energy histories are bit-identical."""
    import wavesim_amd as w
    n, opt, src = builder()
    g = w.WaveSim(n, opt)
    sv = g.solver()
    base = sv.plan
    E1, s1 = g.exec(cases.gpu_source(src))
    # the default plan leaves the choice to the launch's workgroup count
    assert "xcd" not in base
    on = ";".join(item + ("/0" if item.count("/") == 1 else "") + "/xcd" for item in base.split(";"))
    off = on.replace("/xcd", "/noxcd")
    for p in (on, off):
        sv.plan = p
        assert sv.plan == p
        E2, s2 = g.exec(cases.gpu_source(src))
        assert np.array_equal(E1, E2) and np.array_equal(s1.diff_energy, s2.diff_energy)
    sv.plan = base
    assert sv.plan == base


def test_sparse_and_dense_source_uploads_agree(gpu):
    """A source touching few voxels is uploaded as its entries and scattered;
    a large one as a dense field.  Overlapping patches add in patch order from
    zero either way: the solves are bit-identical."""
    import wavesim_amd as w
    n, opt, _ = cases.case_2d_disordered()
    sim = w.WaveSim(n, opt)
    Ny, Nx = sim.N[0], sim.N[1]
    small_pos = [[10, 12, 1, 1], [9, 11, 1, 1], [30, 40, 1, 1]]
    small_val = [np.array([[[[1.0 - 0.5j]]]]), np.full((3, 2, 1, 1), 0.25 + 2j), np.array([[[[-0.0 - 0.0j]]]])]
    # a zero patch over half the grid makes the upload dense
    big_pos = [[1, 1, 1, 1]] + small_pos
    big_val = [np.zeros((Ny, Nx // 2 + 1, 1, 1), np.complex64)] + small_val
    out = []
    for pos, val in ((small_pos, small_val), (big_pos, big_val)):
        sv = sim.solver(1)
        sv.set_source(0, pos, val)
        assert sv.solve(200, 1e-8)
        out.append((sv.field(0), sv.result(0)))
    (E1, r1), (E2, r2) = out
    assert r1.iterations == r2.iterations > 0
    assert np.array_equal(E1, E2) and np.array_equal(r1.diff_energy, r2.diff_energy)


def test_library_starts_from_its_shipped_plan(gpu):
    """A solver created through the C ABI alone (no Python plan lookup) of a
    shape with a tuned plan starts from it; other shapes keep the default."""
    import json
    import wavesim_amd as w
    from wavesim_amd import plans
    from wavesim_amd.solver import MBSSolver
    from wavesim_amd.workloads import WORKLOADS
    nmap, opt, _ = WORKLOADS["homogeneous_2d_256"]()
    sim = w.WaveSim(nmap, opt)
    shipped = json.load(open(plans.SHIPPED))[plans.key(sim.N, 1, len(sim.wiggles))]["plan"]
    with MBSSolver(sim.N, sim.gamma, sim.wiggles, float(sim.epsilon), complex(sim.k02e), sim.roi) as sv:
        assert sv.plan == shipped
    n, opt, _ = cases.case_2d_disordered()
    g = w.WaveSim(n, opt)
    with MBSSolver(g.N, g.gamma, g.wiggles, float(g.epsilon), complex(g.k02e), g.roi) as sv:
        assert "xcd" not in sv.plan and "/512" not in sv.plan


def test_plan_rejects_bad_text(gpu):
    from wavesim_amd.capi import WsimError
    n, opt, _ = cases.case_3d_disordered()
    _, g = _pair(n, opt)
    sv = g.solver()
    for bad in ("q:4,7/0", "y:4,8/0", "y:6,14/0", "y", "y:4,7/0/256/xdc", "y:4,7/0/100", "y:4,7/0/576",
                "y:4,7/0/0/"):
        with pytest.raises(WsimError):
            sv.plan = bad


def test_plan_with_radix_20_needs_a_specialised_kernel(gpu):
    """Radix 20 runs in plan-specialised kernels only: on a shape without one
    the plan is refused with WSIM_ERR_UNSUPPORTED, the previous plan stays in
    force, and the solve still matches the oracle."""
    from oracle import mbs
    from wavesim_amd.capi import WsimError, WSIM_ERR_UNSUPPORTED
    n, opt, src = cases.case_2d_homogeneous(n=64, b=8)
    o, g = _pair(n, opt)
    sv = g.solver()
    assert sv.N[0] == 80
    before = sv.plan
    with pytest.raises(WsimError) as e:
        sv.plan = "y:20,4/0"
    assert e.value.code == WSIM_ERR_UNSUPPORTED
    assert sv.plan == before
    Eo, so = mbs.run(o, cases.oracle_source(src), record_every=True)
    Eg, sg = g.exec(cases.gpu_source(src))
    _check_solve(Eo, so, Eg, sg)


@pytest.mark.parametrize("builder", [cases.case_2d_disordered, cases.case_periodic_2d, cases.case_3d_disordered])
def test_warm_start_matches_oracle(gpu, builder):
    """Warm start from a smooth E0: same iterations and field as the oracle's
    restatement of the warm-start increment recurrence."""
    from oracle import mbs
    import wavesim_amd as w
    n, opt, src = builder()
    o, g = _pair(n, opt)
    rng = np.random.default_rng(11)
    E0 = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)) * 1e-3
    Ez, sz = mbs.run(o, cases.oracle_source(src))
    Eo, so = mbs.run(o, cases.oracle_source(src), E0=E0, ref_energy=sz.diff_energy[0], record_every=True)
    sv = g.solver()
    sv.set_warm_start(0, E0)
    sv.set_reference_energy(0, sz.diff_energy[0])
    Eg, sg = g.exec(cases.gpu_source(src))
    sv.clear_warm_start()
    sv.set_reference_energy(0, 0.0)
    assert sg.iterations == so.it - 1
    assert rel_l2(Eg, Eo) <= TOL
    assert np.allclose(sg.energy_every, so.energy_every, rtol=1e-4, atol=1e-4 * max(so.energy_every))
    # back to the reference start
    Eg2, sg2 = g.exec(cases.gpu_source(src))
    assert sg2.iterations == sz.it - 1 and rel_l2(Eg2, Ez) <= TOL


def test_set_medium_equals_new_solver(gpu):
    """One solver reused for a second medium gives the second medium's results."""
    import wavesim_amd as w
    n1, opt, src = cases.case_2d_disordered()
    n2 = cases.smooth_disorder(n1.shape, 99, absorption=0.02)
    s1 = w.WaveSim(n1, opt)
    s1.exec(cases.gpu_source(src))
    s2 = w.WaveSim(n2, opt)
    s2.adopt_solver(s1)
    E2, st2 = s2.exec(cases.gpu_source(src))
    E2n, st2n = w.WaveSim(n2, opt).exec(cases.gpu_source(src))
    assert st2.iterations == st2n.iterations and np.array_equal(E2, E2n)


def test_warm_start_device_pointer_equals_host(gpu):
    """A warm-start field handed over as device memory (written by another
    stream) gives the same solve as the same field from host memory."""
    import torch
    import wavesim_amd as w
    from wavesim_amd.workloads import spheres_3d
    nmap, opt, src = spheres_3d(96, 16, seed=4)
    sim = w.WaveSim(nmap, opt)
    sv = sim.solver(1)
    p = sim._prepare(src)
    sv.set_source(0, p.positions, p.values)
    rng = np.random.default_rng(3)
    E0 = (rng.standard_normal(sim.N) + 1j * rng.standard_normal(sim.N)).astype(np.complex64) * 1e-3
    out = []
    for dev in (False, True):
        if dev:
            t = torch.from_numpy(np.asfortranarray(E0).ravel(order="F").view(np.float32).copy()).cuda()
            t.mul_(1.0)  # pending work on torch's stream
            sv.set_warm_start(0, device_ptr=int(t.data_ptr()))
        else:
            sv.set_warm_start(0, E0)
        sv.solve(16, 0.0)
        out.append(sv.field(0))
        sv.clear_warm_start()
    assert np.array_equal(out[0], out[1])


def test_failure_exit_on_non_finite_source(gpu):
    """A NaN in the source makes the first increment energy non-finite: the
    solve stops there with WSIM_ERR_DIVERGED instead of running to the cap
    (the reference would run 100000 iterations and report not converged)."""
    import wavesim_amd as w
    from wavesim_amd import capi
    n, opt, _ = cases.case_2d_disordered()
    g = w.WaveSim(n, opt)
    with pytest.raises(capi.WsimError) as ei:
        g.exec(w.Source(np.array([1.0, np.nan, 1.0]), [10, 12]))
    assert ei.value.code == capi.WSIM_ERR_DIVERGED
    r = g.solver().result(0)
    print(f"NaN source: stopped after {r.iterations} iterations", flush=True)
    assert r.diverged and not r.converged and r.iterations == 1


@pytest.mark.parametrize("builder,vector", [(cases.case_2d_disordered, False), (cases.case_3d_disordered, False),
                                            (cases.case_3d_vector_sphere, True)])
def test_failure_exit_on_epsilon_below_the_bound(gpu, builder, vector):
    """epsilon forced to half the convergence bound (opt.epsilon,
    WaveSimBase.m:63): the series diverges and the solve stops with
    WSIM_ERR_DIVERGED where the restated rule stops on the oracle's energies
    (within one wiggle cycle), not at the cap."""
    import wavesim_amd as w
    from oracle import mbs
    from wavesim_amd import capi
    n, opt, src = builder()
    cls = w.WaveSimVector if vector else w.WaveSim
    eps = float(cls(n, opt).epsilon)
    opt2 = dict(opt, epsilon=0.5 * eps)
    o = mbs.make_sim(n, opt2, vector=vector)
    nw = len(o.wiggles)
    with np.errstate(all="ignore"):
        _, so = mbs.run(o, cases.oracle_source(src), record_every=True, max_iterations=300)
    want = cases.divergence_stop(so.energy_every, nw)
    g = cls(n, opt2)
    with pytest.raises(capi.WsimError) as ei:
        g.exec(cases.gpu_source(src))
    assert ei.value.code == capi.WSIM_ERR_DIVERGED
    r = g.solver().result(0)
    print(f"{builder.__name__}: stopped after {r.iterations} iterations (rule on the oracle: {want})", flush=True)
    assert r.diverged and want is not None and abs(r.iterations - want) <= nw


@pytest.mark.parametrize("builder", [cases.case_2d_disordered, cases.case_3d_disordered])
def test_control_once_per_wiggle_cycle_equals_every_iteration(gpu, builder, monkeypatch):
    """Solves run the control step (energy reduction, histories, stop rule)
    once per wiggle cycle, iteration j of a cycle writing its energy partials
    into slot j; WSIM_CYCLE_CONTROL=0 runs it after every iteration.  Same
    iterations, histories and field, bit for bit."""
    import wavesim_amd as w
    n, opt, src = builder()
    out = []
    for mode in ("1", "0"):
        monkeypatch.setenv("WSIM_CYCLE_CONTROL", mode)
        g = w.WaveSim(n, opt)
        E, st = g.exec(cases.gpu_source(src))
        out.append((E, st))
    (E1, s1), (E0, s0) = out
    assert s1.iterations == s0.iterations
    assert np.array_equal(s1.diff_energy, s0.diff_energy)
    assert np.array_equal(s1.energy_every, s0.energy_every)
    assert np.array_equal(E1, E0)
