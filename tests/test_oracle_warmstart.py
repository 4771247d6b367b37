"""Pins of the warm-start preprocessing oracle (oracle/warmstart.py) and of the
host-side prior, by closed forms and special cases."""
import numpy as np

from oracle import analytic
from oracle import warmstart as O


def test_prior_green_matches_reference_utility_restatement():
    from wavesim_amd.learned.prior import green_1d
    k, h = 2 * np.pi * 1.33, 0.25
    z = np.arange(-40, 41) * h
    assert np.allclose(green_1d(k, h, z), analytic.homogeneous_1d(k, h, z), rtol=1e-10, atol=1e-14)


def test_wkb_homogeneous_is_the_background():
    u = np.exp(0.3j * np.arange(8))
    n = np.full((8, 4, 5), 1.33)
    phase, P = O.wkb_prior(n, u, 1.7, 1.33)
    assert np.all(phase == 0) and np.allclose(P, u[:, None, None] * np.ones((1, 4, 5)))


def test_wkb_constant_contrast_is_linear_phase():
    """n - n_bg = d everywhere: phi(z) = k0 h d (z + 1/2) (midpoint rule)."""
    d, k0h = 0.05, 1.3
    n = np.full((10, 3, 2), 1.33 + d)
    phase, P = O.wkb_prior(n, np.ones(10), k0h, 1.33)
    z = np.arange(10)[:, None, None]
    assert np.allclose(phase, k0h * d * (z + 0.5) * np.ones((1, 3, 2)), atol=1e-12)
    assert np.allclose(np.abs(P), 1.0)


def test_wkb_phase_is_path_integral_per_ray():
    """A slab of contrast d from z = a to b: beyond it every ray has picked up
    k0 h d (b - a), before it nothing."""
    n = np.full((12, 2, 2), 1.0)
    n[4:7] += 0.2
    phase, _ = O.wkb_prior(n, np.ones(12), 2.0, 1.0)
    assert np.allclose(phase[:4], 0.0) and np.allclose(phase[7:], 2.0 * 0.2 * 3, atol=1e-12)


def test_features_pool_and_coordinates():
    rng = np.random.default_rng(0)
    n = 1.33 + 0.1 * rng.random((8, 8, 8))
    phase = rng.random((8, 8, 8))
    X = O.features(n, phase, 2, 1.33)
    assert X.shape == (6, 4, 4, 4)
    # block mean of the contrast, by explicit loops
    c = n ** 2 - 1.33 ** 2
    assert np.isclose(X[0, 1, 2, 3], c[2:4, 4:6, 6:8].mean())
    assert np.isclose(X[1, 0, 0, 0], np.cos(phase[:2, :2, :2]).mean())
    # coordinates: cell centres of [-1/2, 1/2], zero mean, along z / y / x
    assert np.allclose(X[3].mean(), 0) and np.allclose(X[3, :, 0, 0], (np.arange(4) + 0.5) / 4 - 0.5)
    assert np.allclose(X[5, 0, 0, :], (np.arange(4) + 0.5) / 4 - 0.5)


def test_compose_zero_envelope_is_prior():
    grid, lo, roi, foot = (10, 9, 8), (1, 2, 1), (4, 4, 6), (1, 5, 2, 6)
    rng = np.random.default_rng(1)
    P = rng.standard_normal((6, 4, 4)) + 1j * rng.standard_normal((6, 4, 4))
    u = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    E = O.compose(grid, lo, roi, 2, P, np.zeros((2, 3, 2, 2)), u, foot)
    assert np.allclose(E[1:5, 2:6, 1:7], P.transpose(1, 2, 0))
    assert np.allclose(E[1:5, 2:6, 0], u[0]) and np.allclose(E[1:5, 2:6, 7], u[7])
    assert np.all(E[0] == 0) and np.all(E[:, 0] == 0)
    # envelope A = (1, 0) doubles the prior in the ROI
    A = np.zeros((2, 3, 2, 2))
    A[0] = 1.0
    E2 = O.compose(grid, lo, roi, 2, P, A, u, foot)
    assert np.allclose(E2[1:5, 2:6, 1:7], 2 * P.transpose(1, 2, 0))
