"""Pins of the Mie-theory oracle (oracle/mie.py) by identities and limits
that fix it independently of its own formulas: energy conservation, the
boundary conditions at the sphere's surface, the Rayleigh dipole and the
electrostatic internal field."""
import numpy as np
import pytest

from oracle import mie


@pytest.mark.parametrize("x,m", [(0.3, 1.2), (3.0, 1.33), (12.5, 1.6 / 1.33), (41.8, 1.2)])
def test_lossless_sphere_conserves_energy(x, m):
    """For real m, Re(a_n) = |a_n|^2 (and for b_n), so Q_ext = Q_sca (BH 4.61,
    4.62): a wrong sign or factor in a_n/b_n breaks it."""
    qext, qsca = mie.efficiencies(x, m)
    assert qext > 0 and abs(qext - qsca) <= 1e-10 * qext


def test_absorbing_sphere_extinguishes_more_than_it_scatters():
    qext, qsca = mie.efficiencies(2.0, 1.5 + 0.1j)
    assert qext > qsca > 0


def test_large_sphere_extinction_paradox():
    """Q_ext -> 2 for x >> 1 (the extinction paradox, BH 4.4.4)."""
    qext, _ = mie.efficiencies(400.0, 1.5 + 0.05j)
    assert abs(qext - 2.0) < 0.05


@pytest.mark.parametrize("m", [1.2, 1.6 / 1.33, 1.5 + 0.02j])
def test_surface_boundary_conditions(m):
    """Across r = a the tangential total E is continuous and the normal D
    (m^2 E_r inside, E_r outside) is continuous: pins the internal
    coefficients, the vector harmonics and the field assembly together."""
    a, lam = 1.0, 1.0
    rng = np.random.default_rng(2)
    th = np.arccos(rng.uniform(-1, 1, 40))
    ph = rng.uniform(-np.pi, np.pi, 40)
    rhat = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=-1)
    eps = 1e-7
    pin, pout = rhat * (a - eps), rhat * (a + eps)
    k = 2 * np.pi / lam
    inc = lambda p: np.stack([np.exp(1j * k * p[:, 2]), 0 * p[:, 2], 0 * p[:, 2]], axis=-1)
    Ein = mie.near_field(a, m, 1.0, lam, pin[:, 0], pin[:, 1], pin[:, 2]) + inc(pin)
    Eout = mie.near_field(a, m, 1.0, lam, pout[:, 0], pout[:, 1], pout[:, 2]) + inc(pout)
    tin = Ein - np.sum(Ein * rhat, axis=1, keepdims=True) * rhat
    tout = Eout - np.sum(Eout * rhat, axis=1, keepdims=True) * rhat
    scale = np.abs(Eout).max()
    assert np.abs(tin - tout).max() <= 1e-5 * scale
    assert np.abs(m ** 2 * np.sum(Ein * rhat, axis=1) - np.sum(Eout * rhat, axis=1)).max() <= 1e-5 * scale


def test_rayleigh_sphere_is_a_dipole():
    """x << 1: the scattered near field is that of the dipole p = a^3
    (m^2-1)/(m^2+2) x_hat (Gaussian units): e^{ikr}[k^2 (r x p) x r / r +
    (3 r (r.p) - p)(1/r^3 - i k/r^2)] (Jackson 9.18)."""
    a, lam, m = 0.002, 1.0, 1.5
    k = 2 * np.pi / lam
    rng = np.random.default_rng(4)
    d = rng.normal(size=(30, 3))
    rhat = d / np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(3 * a, 0.3 * lam, size=(30, 1))
    pts = rhat * r
    E = mie.near_field(a, m, 1.0, lam, pts[:, 0], pts[:, 1], pts[:, 2])
    p = np.array([1.0, 0.0, 0.0]) * a ** 3 * (m ** 2 - 1) / (m ** 2 + 2)
    rp = (rhat @ p)[:, None]
    far = k ** 2 * (p[None, :] - rhat * rp) / r
    near = (3 * rhat * rp - p[None, :]) * (1 / r ** 3 - 1j * k / r ** 2)
    ref = np.exp(1j * k * r) * (far + near)
    assert np.linalg.norm(E - ref) <= 2e-3 * np.linalg.norm(ref)


def test_small_sphere_internal_field_is_electrostatic():
    """x << 1: the internal field is uniform, 3/(m^2+2) E0 along x (BH 5.2):
    at the centre (where internal minus incident is 3/(m^2+2) - 1) to the
    O(x^2) size correction, elsewhere inside to O(m k r) (its phase varies
    across the sphere, m k a = 0.01 here)."""
    a, m = 0.001, 1.6
    want = 3 / (m ** 2 + 2)
    E0 = mie.near_field(a, m, 1.0, 1.0, np.array([0.0]), np.array([0.0]), np.array([0.0]))[0]
    assert abs(E0[0] - (want - 1)) <= 1e-4 * abs(want - 1)
    assert np.abs(E0[1:]).max() <= 1e-8
    rng = np.random.default_rng(5)
    p = rng.uniform(-0.5 * a, 0.5 * a, size=(20, 3))
    E = mie.near_field(a, m, 1.0, 1.0, p[:, 0], p[:, 1], p[:, 2])
    E[:, 0] += np.exp(2j * np.pi * p[:, 2])
    assert np.abs(E[:, 0] - want).max() <= 1e-2 * want
    assert np.abs(E[:, 1:]).max() <= 1e-2 * want


def test_no_contrast_no_scattering():
    X, Y, Z, _ = mie.reference_grid([8, 6, 5], 0.2)
    E = mie.near_field(1.0, 1.0, 1.0, 1.0, X, Y, Z)
    assert np.abs(E).max() <= 1e-5  # the truncated expansion of the incident wave inside


def test_reference_grid_follows_create_sphere():
    """create_sphere.m:33-37: x has N(2) points offset by N(1)/2 pixels, y has
    N(1) points offset by N(2)/2 pixels."""
    X, Y, Z, (xr, yr, zr) = mie.reference_grid([60, 40, 30], 0.2)
    assert X.shape == (60, 40, 30)
    assert np.allclose(xr, np.arange(1, 41) * 0.2 - 6.0)
    assert np.allclose(yr, np.arange(1, 61) * 0.2 - 4.0)
    assert np.allclose(zr, np.arange(1, 31) * 0.2 - 3.0)
    n = mie.sphere_index([60, 40, 30], 0.2, 1.0, 1.2, 1.0)
    assert n.shape == (60, 40, 30) and set(np.unique(n)) == {1.0, 1.2}
