"""Pins for the oracle's restatement of the reference set-up code
(SimulationGrid, Medium, ARL filters, epsilon/gamma/k02e, wiggles).

Every expected value here comes from a hand table, a closed form or a library
routine -- not from re-typing the oracle's formula.
"""
import math

import numpy as np
import pytest
import scipy.signal

from oracle import mbs


def test_factor_matches_matlab_convention():
    assert mbs.factor(1) == [1]
    assert mbs.factor(12) == [2, 2, 3]
    assert mbs.factor(2187) == [3] * 7
    assert mbs.factor(13) == [13]
    assert mbs.factor(1100) == [2, 2, 5, 5, 11]


@pytest.mark.parametrize("s,expect", [
    (320, 320), (1152, 1152), (832, 840), (1088, 1100), (2176, 2187),  # config grids
    (13, 14), (17, 18), (19, 20), (23, 24), (26, 27), (29, 30), (31, 32),
    (49, 49), (77, 77), (97, 99), (1, 1), (2, 2)])
def test_efficient_size_hand_table(s, expect):
    # 97 prime; 98 = 2*7*7 has a second-largest factor 7 > 5; 99 = 3*3*11 passes.
    # 49 = 7*7 and 77 = 7*11 have only two factors, so the second rule is off.
    assert mbs.efficient_size(s) == expect


@pytest.mark.parametrize("n", [2, 4, 8, 320, 1152])
def test_symrange_even_is_fft_order(n):
    assert np.array_equal(mbs.symrange(n), np.rint(np.fft.fftfreq(n) * n))


def test_symrange_hand_table():
    # MATLAB fftshift = circshift(x, floor(N/2)); for odd N this is NOT the
    # FFT ordering (that would be ifftshift): the reference assigns frequency
    # +1 to index 0.  The oracle reproduces the reference.
    assert list(mbs.symrange(1)) == [0]
    assert list(mbs.symrange(3)) == [1, -1, 0]
    assert list(mbs.symrange(4)) == [0, 1, -2, -1]
    assert list(mbs.symrange(5)) == [1, 2, -2, -1, 0]


@pytest.mark.parametrize("L", list(range(1, 41)) + [63, 127])
def test_nuttallwin_matches_scipy(L):
    ref = scipy.signal.windows.nuttall(L, sym=True)
    assert np.allclose(mbs.nuttallwin(L), ref, atol=1e-12, rtol=0)


def test_nuttall_published_coefficients():
    # w(0) = a0 - a1 + a2 - a3, centre = a0 + a1 + a2 + a3 = 1
    w = mbs.nuttallwin(5)
    assert w[0] == pytest.approx(0.3635819 - 0.4891775 + 0.1365995 - 0.0106411, abs=1e-15)
    assert w[2] == pytest.approx(1.0, abs=1e-12)
    assert w[1] == pytest.approx(0.3635819 - 0.1365995, abs=1e-12)


def test_medium_padding_and_widths():
    rng = np.random.default_rng(0)
    n = 1.0 + rng.random((7, 6))
    med = mbs.medium(n, dict(pixel_size=0.25, boundary_widths=[3, 2]))
    # y: 7 + 6 = 13 -> 14 ; x: 6 + 4 = 10 ; z periodic 1
    assert med.grid.N == [14, 10, 1]
    assert med.Bl == [4, 2, 0] and med.Br == [3, 2, 0]
    ref = np.pad((n ** 2).reshape(7, 6, 1), [(4, 3), (2, 2), (0, 0)], mode="edge")
    assert np.array_equal(med.e_r, ref)
    assert med.e_r_min == pytest.approx((n ** 2).min())
    assert med.e_r_center == pytest.approx(((n ** 2).min() + (n ** 2).max()) / 2)


def test_medium_rejects_gain():
    with pytest.raises(ValueError):
        mbs.medium(np.array([[1.0 - 0.1j]]), dict(pixel_size=1, boundary_widths=[0, 0]))


def test_homogeneous_constants_closed_form():
    nval = 1.2 + 0.01j
    lam = 1.0
    opt = dict(pixel_size=0.25, boundary_widths=[0, 0], single_precision=False)
    sim = mbs.make_sim(np.full((8, 10), nval), opt)
    k0 = 2 * math.pi / lam
    er = nval ** 2
    k0c = math.sqrt(er.real) * k0
    eps = max(abs(er.imag) * k0 ** 2 / 0.95, 3.0)
    assert sim.epsilon == pytest.approx(eps, rel=1e-12)
    assert sim.k0c == pytest.approx(k0c, rel=1e-12)
    # gamma = i V / eps with V = i Im(e_r) k0^2 - i eps  ->  1 - Im(e_r) k0^2 / eps (real)
    assert np.allclose(sim.gamma, 1 - er.imag * k0 ** 2 / eps, atol=1e-12)
    assert sim.k02e == pytest.approx(k0c ** 2 / eps + 1j)
    assert sim.iterations_per_cycle == pytest.approx(lam / (2 * k0c / eps))
    # periodic everywhere: no filters, one (zero) wiggle
    assert len(sim.wiggles) == 1 and all(f is None for f in sim.filters)


def test_epsilon_minimum():
    sim = mbs.make_sim(np.ones((8, 8)), dict(pixel_size=0.25, boundary_widths=[2, 2]))
    assert sim.epsilon == pytest.approx(3.0)


def test_arl_filter_2d_hand_values():
    # 2-D, L = 3, even padding: [0 ; nuttall(5)(1:2)] then ones, mirrored
    opt = dict(pixel_size=0.25, boundary_widths=[3, 3])
    sim = mbs.make_sim(np.ones((6, 6)), opt)
    a = [0.3635819, 0.4891775, 0.1365995, 0.0106411]
    w0 = a[0] - a[1] + a[2] - a[3]
    w1 = a[0] - a[2]
    exp = np.array([0, w0, w1, 1, 1, 1, 1, 1, 1, w1, w0, 0])
    assert sim.medium.grid.N[:2] == [12, 12]
    assert np.allclose(sim.filters[0], exp, atol=1e-12)
    assert np.allclose(sim.filters[1], exp, atol=1e-12)


def test_arl_filter_odd_padding():
    # y: 7 + 6 = 13 -> 14, Bl = 4 != Br = 3: smoothstep = nuttall(5)(1:3), one leading zero
    opt = dict(pixel_size=0.25, boundary_widths=[3, 3])
    sim = mbs.make_sim(np.ones((7, 6)), opt)
    a = [0.3635819, 0.4891775, 0.1365995, 0.0106411]
    w0 = a[0] - a[1] + a[2] - a[3]
    w1 = a[0] - a[2]
    exp = np.array([0, w0, w1, 1.0] + [1] * 7 + [1.0, w1, w0])
    assert np.allclose(sim.filters[0], exp, atol=1e-12)


def test_arl_filter_3d_linear_window():
    opt = dict(pixel_size=0.25, boundary_widths=[2, 2, 2])
    sim = mbs.make_sim(np.ones((4, 4, 4)), opt)
    s = 0.79 / 1.66  # ((1) - 0.21) / (1 + 0.66)
    exp = np.array([0, s, 1, 1, 1, 1, s, 0])
    for d in range(3):
        assert np.allclose(sim.filters[d], exp, atol=1e-12)


def test_wiggle_perm_hand_tables():
    assert np.array_equal(mbs.wiggle_perm([True, False, False]), [[1, -1], [0, 0], [0, 0]])
    t2 = mbs.wiggle_perm([True, True, False])
    assert np.array_equal(t2.T, [[1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]])
    t3 = mbs.wiggle_perm([True, True, True])
    assert np.array_equal(t3.T, [[1, 1, 1], [-1, 1, 1], [1, -1, 1], [1, 1, -1],
                                 [-1, -1, 1], [-1, 1, -1], [1, -1, -1], [-1, -1, -1]])
    tz = mbs.wiggle_perm([True, False, True])
    assert np.array_equal(tz.T, [[1, 0, 1], [-1, 0, 1], [-1, 0, -1], [1, 0, -1]])
    assert mbs.wiggle_perm([False, False, False]).shape == (3, 1)


def test_wiggle_descriptor_closed_form():
    opt = dict(pixel_size=0.5, boundary_widths=[4, 4], single_precision=False)
    sim = mbs.make_sim(np.ones((10, 12)), opt)
    g = sim.medium.grid
    assert len(sim.wiggles) == 4
    for w, wig in enumerate(sim.wiggles):
        sy, sx = sim.wiggle_set[0, w], sim.wiggle_set[1, w]
        n = np.arange(g.N[1])
        assert np.allclose(wig["gx"], np.exp(1j * np.pi / 2 * sx * n / g.N[1]))
        assert np.allclose(wig["gy"], np.exp(1j * np.pi / 2 * sy * np.arange(g.N[0]) / g.N[0]))
        px = 2 * np.pi / (0.5 * g.N[1]) * np.fft.fftfreq(g.N[1]) * g.N[1]
        assert np.allclose(wig["pxe"], (px - sx * 2 * np.pi / (0.5 * g.N[1]) / 4) / np.sqrt(sim.epsilon))
        assert np.allclose(wig["gz"], 1) and np.allclose(wig["pze"], 0)


def test_roi_defaults_and_vector():
    opt = dict(pixel_size=0.25, boundary_widths=[2, 3])
    sim = mbs.make_sim(np.ones((6, 8)), opt)
    Bl = sim.medium.Bl
    assert sim.roi.tolist() == [[Bl[0] + 1, Bl[1] + 1, 1, 1],
                                [sim.N[0] - sim.medium.Br[0], sim.N[1] - sim.medium.Br[1], 1, 1]]
    vs = mbs.make_sim(np.ones((6, 8, 4)), dict(pixel_size=0.25, boundary_widths=[1, 1, 1]), vector=True)
    assert vs.N[3] == 3 and vs.roi[1, 3] == 3 and vs.roi[0, 3] == 1
    us = mbs.make_sim(np.ones((6, 8)), dict(opt, roi=[[2, 3], [4, 5]]))
    assert us.roi[0, :2].tolist() == [Bl[0] + 2, Bl[1] + 3]
    assert us.roi[1, :2].tolist() == [Bl[0] + 4, Bl[1] + 5]


def test_source_shift_crop():
    s = mbs.Source.point(np.arange(6.0).reshape(3, 2), [0, 4])
    c = s.shift([2, 2, 1, 1]).crop(np.array([[1, 1, 1, 1], [2, 5, 1, 1]]))
    # original corner (0,4) -> (1,5); rows 1..3 and cols 5..6 cropped to rows 1..2, col 5
    assert c.positions == [[1, 5, 1, 1]]
    assert np.array_equal(c.values[0][..., 0, 0], [[0.0], [2.0]])
    assert mbs.Source.point(1.0, [9, 9]).shift([1, 1, 1, 1]).crop(
        np.array([[1, 1, 1, 1], [4, 4, 1, 1]])).positions == []
