"""Pins of the PSTD oracle (oracle/pstd.py, a restatement of PSTD.m) by what
the mathematics fixes, independent of the restatement's own formulas."""
import math

import numpy as np
import pytest

from oracle import mbs, pstd


def _opt(**kw):
    o = dict(pixel_size=0.25, boundary_widths=[0, 0], boundary_type="ARL", energy_threshold=1e-8,
             callback_interval=50, single_precision=False)
    o.update(kw)
    return o


def test_options_boundary_type_required_and_strength_refused():
    """PSTD.m:32-34: boundary_type must be given, boundary_strength must not."""
    n = np.ones((8, 8))
    with pytest.raises(ValueError):
        pstd.make_pstd(n, dict(pixel_size=0.25, boundary_widths=[0, 0]))
    with pytest.raises(ValueError):
        pstd.make_pstd(n, dict(pixel_size=0.25, boundary_widths=[0, 0], boundary_type="ARL",
                               boundary_strength=0.2))
    pstd.make_pstd(n, _opt())


@pytest.mark.parametrize("shape,dims", [((16, 12), 2), ((8, 10, 6), 3), ((16, 1), 2)])
def test_time_step_and_lossless_coefficients_by_hand(shape, dims):
    """dtmax = 2 dx sqrt(e_r,min) / (pi sqrt(dims)), dt = 0.95 dtmax, one cycle
    = lambda / dt steps; without absorption c1 = -1, c2 = 2, c3 = dt^2 / e_r
    (the leapfrog E+ = 2E - E- + dt^2/e_r (lap E + S))."""
    n = np.full(shape, 1.5)
    o = pstd.make_pstd(n, _opt(boundary_widths=[0] * len(shape), **{"lambda": 0.8}))
    dtmax = 2 * 0.25 * 1.5 / (math.pi * math.sqrt(dims))
    assert o.dimensions == dims
    assert o.dtmax == pytest.approx(dtmax, rel=1e-14)
    assert o.dt == pytest.approx(0.95 * dtmax, rel=1e-14)
    assert o.iterations_per_cycle == pytest.approx(0.8 / (0.95 * dtmax), rel=1e-14)
    assert np.allclose(o.c1, -1.0, rtol=0, atol=1e-15)
    assert np.allclose(o.c2, 2.0, rtol=0, atol=1e-15)
    assert np.allclose(o.c3, o.dt ** 2 / 2.25, rtol=1e-14)


def test_laplacian_of_grid_plane_waves():
    """ifftn(koperator fftn(E)) = -|p|^2 E for E = exp(i p.r) with p on the grid."""
    o = pstd.make_pstd(np.ones((12, 10, 8)), _opt(boundary_widths=[0, 0, 0]))
    dx = 0.25
    y, x, z = np.meshgrid(np.arange(12) * dx, np.arange(10) * dx, np.arange(8) * dx, indexing="ij")
    for ky, kx, kz in [(1, 0, 0), (0, 3, 0), (2, -4, 3), (6, 5, -4)]:
        p = (2 * math.pi * ky / (12 * dx), 2 * math.pi * kx / (10 * dx), 2 * math.pi * kz / (8 * dx))
        E = np.exp(1j * (p[0] * y + p[1] * x + p[2] * z))[..., None]
        L = pstd.laplacian(o, E)
        assert np.allclose(L, -(p[0] ** 2 + p[1] ** 2 + p[2] ** 2) * E, rtol=0, atol=1e-9 * (p[0] ** 2 + p[1] ** 2 + p[2] ** 2 + 1))


def test_lossless_mode_follows_cos_k_theta():
    """One Fourier mode without source in a lossless medium: with E_0 = m and
    E_-1 = cos(theta) m the recurrence gives E_k = cos(k theta) m,
    cos(theta) = 1 - dt^2 |p|^2 / (2 e_r)."""
    o = pstd.make_pstd(np.full((16, 20), 1.2), _opt())
    dx = 0.25
    y, x = np.meshgrid(np.arange(16) * dx, np.arange(20) * dx, indexing="ij")
    p = (2 * math.pi * 3 / (16 * dx), 2 * math.pi * -2 / (20 * dx))
    m = np.exp(1j * (p[0] * y + p[1] * x))[..., None, None]
    cth = 1 - o.dt ** 2 * (p[0] ** 2 + p[1] ** 2) / (2 * 1.44)
    th = math.acos(cth)
    Em, E = cth * m, m
    for k in range(1, 41):
        Em, E = E, pstd.step(o, E, Em, None, 0.0)
        assert np.allclose(E, math.cos(k * th) * m, rtol=0, atol=1e-10)


def test_converged_field_is_the_time_harmonic_solution_of_the_recurrence():
    """Homogeneous absorbing periodic medium: after the source has ramped up
    and transients have decayed, the phase-compensated output equals the
    closed-form time-harmonic solution X / z of the recurrence (apart from the
    static p = 0 offset, which the damped wave equation does not remove)."""
    n = np.full((24, 32), 1.0 + 0.05j)
    o = pstd.make_pstd(n, _opt(energy_threshold=0.0))
    S = mbs.Source.point(1.0, [12, 16])
    E, st = pstd.run(o, S, max_iterations=3000)
    assert not st.converged and st.it - 1 >= 3000
    X = pstd.steady_state_periodic(o, S.dense(o.N, np.complex128))
    a, b = E - E.mean(), X - X.mean()
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-8


def test_discrete_steady_state_tends_to_helmholtz_second_order():
    """The time-harmonic solution of the recurrence approaches the Helmholtz
    solution S / (|p|^2 - k0^2 e_r) as dt -> 0, with second-order error: this
    fixes c1, c2, c3 (absorption sign and scale) and the Laplacian's sign."""
    n = np.full((24, 32), 1.0 + 0.05j)
    errs = []
    for dtr in (0.2, 0.05):
        o = pstd.make_pstd(n, _opt(dt_relative=dtr))
        Sd = mbs.Source.point(1.0, [12, 16]).dense(o.N, np.complex128)
        X = pstd.steady_state_periodic(o, Sd)[..., 0]
        H = np.fft.ifftn(np.fft.fftn(Sd[..., 0]) / (-o.koperator - (2 * math.pi) ** 2 * o.medium.e_r))
        errs.append(np.linalg.norm(X - H) / np.linalg.norm(H))
    assert errs[1] < 1.5e-3
    assert 12 < errs[0] / errs[1] < 20  # (0.2 / 0.05)^2 = 16


def test_default_source_ramp():
    """PSTD.m:130-137: amplitude 0 at N = 0, half at 50 cycles, 1 from 100 on,
    phase exp(-2 pi i N)."""
    assert pstd.default_source(0.0) == 0
    assert abs(pstd.default_source(50.0)) == pytest.approx(0.5, abs=1e-12)
    assert pstd.default_source(100.0) == pytest.approx(1.0, abs=1e-9)
    assert pstd.default_source(150.25) == pytest.approx(-1j, abs=1e-9)


def test_stop_rule_waits_for_real_part_above_0_9():
    """Simulation.next with PSTD's can_terminate = A > 0.9 (MATLAB compares the
    real part): a stop happens only on a step whose source amplitude has real
    part above 0.9, i.e. after about 80 cycles and near phase 0."""
    n = np.full((16, 16), 1.0 + 0.1j)
    o = pstd.make_pstd(n, _opt(energy_threshold=1e-3))
    E, st = pstd.run(o, mbs.Source.point(1.0, [8, 8]))
    K = st.it - 1
    A = pstd.default_source(K / o.iterations_per_cycle)
    assert st.converged and A.real > 0.9
    assert K / o.iterations_per_cycle > 79.5
    # earlier steps that met the threshold could not stop
    de = np.asarray(st.diff_energy)
    early = [k for k in range(1, K) if de[k - 1] / de[0] < 1e-3]
    assert early and all(pstd.default_source(k / o.iterations_per_cycle).real <= 0.9 for k in early)
