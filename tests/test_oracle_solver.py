"""Pins for the oracle's restatement of the MBS loop: closed-form solutions,
brute-force operators, invariants and hand-traced stopping decisions."""
import math

import numpy as np
import pytest

from oracle import analytic, mbs
from tests import cases
from tests.cases import rel_l2


def _dft(n, sign=-1):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def test_propagate_bruteforce_dense_dft():
    """propagate() against explicit DFT matrices on a tiny non-square grid
    (pins fftn sign/normalisation, axis order and the broadcast of pxe/pye)."""
    opt = dict(pixel_size=0.3, boundary_widths=[2, 1], single_precision=False)
    sim = mbs.make_sim(np.full((4, 7), 1.1 + 0.05j), opt)
    Ny, Nx = sim.N[:2]
    assert Ny != Nx and len(sim.wiggles) == 4
    rng = np.random.default_rng(3)
    E = rng.standard_normal((Ny, Nx, 1, 1)) + 1j * rng.standard_normal((Ny, Nx, 1, 1))
    for wig in sim.wiggles:
        g = np.outer(wig["gy"], wig["gx"])
        A = E[:, :, 0, 0] * g
        F = _dft(Ny) @ A @ _dft(Nx).T
        den = wig["pye"][:, None] ** 2 + wig["pxe"][None, :] ** 2 + wig["pze"][0] ** 2 - sim.k02e
        B = (_dft(Ny, +1) @ (F / den) @ _dft(Nx, +1).T) / (Ny * Nx)
        ref = B * np.conj(g)
        out = mbs.propagate(sim, E, wig)[:, :, 0, 0]
        assert rel_l2(out, ref) < 1e-12


def test_mix_closed_form():
    g = np.array([0.3 + 0.2j, 1.0, 0.0])
    dE = np.array([1.0 + 1j, 2.0, 3.0 - 1j])
    P = np.array([0.5j, -1.0, 2.0])
    out = mbs.f_mix(dE, P, g, single=False)
    exp = (1 - 0.7 * g) * dE - 1j * 0.7 * g * g * P
    assert np.allclose(out, exp)
    assert out[2] == dE[2]  # gamma = 0 leaves dE untouched


def test_vector_green_invariants():
    rng = np.random.default_rng(4)
    f = [rng.standard_normal(50) + 1j * rng.standard_normal(50) for _ in range(3)]
    p = [rng.standard_normal(50) for _ in range(3)]
    k = 2.3 + 1j
    Gx, Gy, Gz = mbs.f_g0_vector(*f, *p, k)
    pf = p[0] * f[0] + p[1] * f[1] + p[2] * f[2]
    # p . G f = -(p . f) / k02e
    assert np.allclose(p[0] * Gx + p[1] * Gy + p[2] * Gz, -pf / k)
    # transverse part is the scalar Green's function
    g0 = 1 / (p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - k)
    pp = p[0] ** 2 + p[1] ** 2 + p[2] ** 2
    ft = [fc - pc * pf / pp for fc, pc in zip(f, p)]
    Gt = mbs.f_g0_vector(*ft, *p, k)
    for a, b in zip(Gt, ft):
        assert np.allclose(a, g0 * b)


def test_periodic_homogeneous_closed_form():
    nval = 1.0 + 0.05j
    sim = mbs.make_sim(np.full((32, 24), nval), dict(pixel_size=0.25, boundary_widths=[0, 0],
                                                     energy_threshold=1e-16, single_precision=False))
    E, st = mbs.run(sim, mbs.Source.point(1.0, [10, 7]))
    assert st.converged
    S = np.zeros((32, 24, 1))
    S[9, 6, 0] = 1
    ref = analytic.periodic_homogeneous(S, sim.k0, nval, 0.25)
    assert rel_l2(E[..., 0], ref) < 1e-7


def test_periodic_disordered_helmholtz_fixed_point():
    n, opt, src = cases.case_periodic_2d(thr=1e-16)
    opt = dict(opt, single_precision=False)
    sim = mbs.make_sim(n, opt)
    E, st = mbs.run(sim, cases.oracle_source(src))
    S = mbs.prepare_source(sim, cases.oracle_source(src)).dense(sim.N, np.complex128)[..., 0]
    res = analytic.helmholtz_residual(E[..., 0], S, sim.k0, n.reshape(n.shape + (1,)), 0.25)
    assert np.linalg.norm(res) / np.linalg.norm(S) < 1e-6


def test_single_vs_double_precision():
    n, opt, src = cases.case_2d_disordered()
    s1 = mbs.make_sim(n, opt)
    s2 = mbs.make_sim(n, dict(opt, single_precision=False))
    E1, st1 = mbs.run(s1, cases.oracle_source(src))
    E2, st2 = mbs.run(s2, cases.oracle_source(src), max_iterations=None)
    assert E1.dtype == np.complex64 and E2.dtype == np.complex128
    assert abs((st1.it - st2.it)) <= len(s1.wiggles)
    k = min(st1.it, st2.it) - 1
    E1k, _ = mbs.run(s1, cases.oracle_source(src), max_iterations=k)
    s2b = mbs.make_sim(n, dict(opt, single_precision=False, energy_threshold=0.0))
    E2k, _ = mbs.run(s2b, cases.oracle_source(src), max_iterations=k)
    assert rel_l2(E1k, E2k) < 1e-5


def test_linearity_in_source():
    n, opt, _ = cases.case_2d_disordered()
    sim = mbs.make_sim(n, dict(opt, energy_threshold=0.0, single_precision=False))
    a = mbs.Source.point(1.0, [10, 12])
    b = mbs.Source.point(0.5j, [25, 30])
    Ea, _ = mbs.run(sim, a, max_iterations=24)
    Eb, _ = mbs.run(sim, b, max_iterations=24)
    Eab, _ = mbs.run(sim, a + b, max_iterations=24)
    assert rel_l2(Eab, Ea + Eb) < 1e-12


def test_1d_analytic_solution():
    """ARL + ACC in 1-D against homogeneous_medium_analytic_solution."""
    Ny = 256
    sim = mbs.make_sim(np.ones((Ny, 1)), dict(pixel_size=0.25, boundary_widths=[32, 0],
                                             energy_threshold=1e-12, single_precision=False))
    E, st = mbs.run(sim, mbs.Source.point(1.0, [Ny // 2, 1]))
    x = (np.arange(Ny) - (Ny // 2 - 1)) * 0.25
    ref = analytic.homogeneous_1d(sim.k0, 0.25, x)
    assert rel_l2(E[:, 0, 0, 0], ref) < 1e-4


def test_analytic_1d_limit_value():
    # the closed form at x = 0 equals the limit of the general expression
    k, h = 2 * math.pi, 0.25
    v0 = analytic.homogeneous_1d(k, h, [0.0])[0]
    v1 = analytic.homogeneous_1d(k, h, [1e-7])[0]
    assert abs(v0 - v1) < 1e-5 * abs(v0)


@pytest.mark.slow
def test_2d_hankel_far_field():
    """Point source in 2-D free space (ARL + ACC) vs (i h^2/4) H0(k r) in an annulus."""
    n = 128
    sim = mbs.make_sim(np.ones((n, n)), dict(pixel_size=0.25, boundary_widths=[32, 32],
                                            energy_threshold=1e-10, single_precision=False))
    c = n // 2
    E, st = mbs.run(sim, mbs.Source.point(1.0, [c, c]))
    yy, xx = np.meshgrid(np.arange(n) - (c - 1), np.arange(n) - (c - 1), indexing="ij")
    r = np.hypot(yy, xx) * 0.25
    ref = analytic.hankel_2d(sim.k0, 0.25, np.where(r > 0, r, 1))
    mask = (r > 2.0) & (r < 14.0)
    assert rel_l2(E[..., 0, 0][mask], ref[mask]) < 2e-2


def test_stopping_rule_hand_traced():
    # energies 0.5^(it-1); energy refreshed at it = 1, 2, 10, 18, ... ; tests at multiples of Nw = 4
    e = [0.5 ** k for k in range(200)]
    it, conv, de = mbs.stop_iteration(e, 4, 0.01, 8)
    assert (it, conv) == (12, True)          # de(12) = e(10) = 2^-9 < 0.01
    assert de[2:9] == [0.5] * 7              # stale value of iteration 2
    it, conv, _ = mbs.stop_iteration(e, 4, 1e-3, 8)
    assert (it, conv) == (20, True)          # de(16) = 2^-9 > 1e-3, de(20) = e(18) = 2^-17
    it, conv, _ = mbs.stop_iteration(e, 4, 0.0, 8, max_iterations=10)
    assert (it, conv) == (12, False)         # the cap is only tested at multiples of Nw
    it, conv, _ = mbs.stop_iteration(e, 1, 0.01, 1)
    assert (it, conv) == (8, True)           # 0.5^7 < 0.01 <= 0.5^6


def test_run_follows_stopping_rule():
    n, opt, src = cases.case_2d_disordered()
    sim = mbs.make_sim(n, opt)
    E, st = mbs.run(sim, cases.oracle_source(src), record_every=True)
    it, conv, de = mbs.stop_iteration(st.energy_every, len(sim.wiggles), sim.energy_threshold,
                                      sim.energy_calculation_interval)
    assert it == st.it - 1 and conv == st.converged
    assert np.allclose(de, st.diff_energy)


def test_zero_source_returns_early():
    sim = mbs.make_sim(np.ones((8, 8)), dict(pixel_size=0.25, boundary_widths=[2, 2]))
    E, st = mbs.run(sim, mbs.Source.point(0.0, [3, 3]))
    assert E is None and st.source_energy == 0


def test_source_outside_grid_is_cropped():
    sim = mbs.make_sim(np.ones((8, 8)), dict(pixel_size=0.25, boundary_widths=[2, 2]))
    s = mbs.prepare_source(sim, mbs.Source.point(np.ones((4, 1)), [-5, 3]))
    assert s.positions == []


def test_warm_start_zero_equals_zero_start():
    """E0 = 0 is exactly the reference loop, with and without wiggles."""
    for builder in (cases.case_periodic_2d, cases.case_2d_disordered):
        n, opt, src = builder()
        sim = mbs.make_sim(n, dict(opt, single_precision=False))
        E1, s1 = mbs.run(sim, cases.oracle_source(src))
        E2, s2 = mbs.run(sim, cases.oracle_source(src), E0=np.zeros(sim.N))
        assert s1.it == s2.it and rel_l2(E2, E1) < 1e-14


def test_warm_start_from_solution_stops_at_once_nw1():
    n, opt, src = cases.case_periodic_2d(thr=1e-14)
    sim = mbs.make_sim(n, dict(opt, single_precision=False))
    E, s = mbs.run(sim, cases.oracle_source(src))
    sim2 = mbs.make_sim(n, dict(opt, single_precision=False, energy_threshold=1e-6))
    Ew, sw = mbs.run(sim2, cases.oracle_source(src), E0=s.E_primed / sim.gamma[..., None],
                     ref_energy=s.diff_energy[0])
    assert sw.it - 1 == 1                       # the first stop test passes
    assert rel_l2(Ew, E) < 1e-6


def test_warm_start_with_wiggles_from_solution():
    """With Nw > 1 the zero-start answer is not the fixed point of one
    wiggle's step; spreading the correction over the wiggle cycle keeps a
    warm start from the converged field close to it and quick."""
    n, opt, src = cases.case_2d_disordered(thr=1e-14)
    sim = mbs.make_sim(n, dict(opt, single_precision=False))
    E, s = mbs.run(sim, cases.oracle_source(src))
    with np.errstate(all="ignore"):
        Efull = s.E_primed / sim.gamma[..., None]
    Efull[~np.isfinite(Efull)] = 0
    sim2 = mbs.make_sim(n, dict(opt, single_precision=False, energy_threshold=1e-6))
    _, sz = mbs.run(sim2, cases.oracle_source(src))
    Ew, sw = mbs.run(sim2, cases.oracle_source(src), E0=Efull, ref_energy=s.diff_energy[0])
    print("zero", sz.it - 1, "warm", sw.it - 1, "err", rel_l2(Ew, E))
    assert sw.converged and sw.it < sz.it and rel_l2(Ew, E) < 2e-3


def test_warm_start_converges_to_fixed_point_nw1():
    n, opt, src = cases.case_periodic_2d(thr=1e-14)
    sim = mbs.make_sim(n, dict(opt, single_precision=False))
    E, s = mbs.run(sim, cases.oracle_source(src))
    rng = np.random.default_rng(0)
    E0 = 1e-3 * (rng.standard_normal(sim.N) + 1j * rng.standard_normal(sim.N))
    Ew, sw = mbs.run(sim, cases.oracle_source(src), E0=E0, ref_energy=s.diff_energy[0])
    assert rel_l2(Ew, E) < 1e-6
