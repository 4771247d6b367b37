"""The library's C++ problem set-up (wsim_problem_query / wsim_problem_build,
wrapped by wavesim_amd.problem.build) against the Python mirror of the
WaveSim / WaveSimVector constructor, which tests/test_capi.py pins to the
oracle: the same grid, ROI, gamma, epsilon, k02e and wiggle tables.  The
C++ code rounds as numpy does on x86-64 (a fused multiply-add in the real part
of complex products, |z| = max sqrt(fma(r, r, 1))); the checks allow one unit
in the last place where numpy could take another path on another CPU.
Host code only (no GPU)."""
import numpy as np
import pytest

from tests import cases


def _compare(nmap, opt, vector=False):
    import wavesim_amd as w
    from wavesim_amd import problem
    g = (w.WaveSimVector if vector else w.WaveSim)(nmap, opt)
    p = problem.build(nmap, opt, vector=vector)
    assert p["N"] == list(g.N)
    assert np.array_equal(p["roi"], g.roi)
    assert p["Bl"] == list(g.sample.Bl) and p["Br"] == list(g.sample.Br)
    assert p["epsilon"] == g.epsilon and p["k02e"] == g.k02e
    assert abs(p["k0c"] - g.k0c) <= 2e-16 * g.k0c
    assert abs(p["iterations_per_cycle"] - g.iterations_per_cycle) <= 4e-16 * g.iterations_per_cycle
    gp = np.ascontiguousarray(p["gamma"].ravel(order="F")).view(np.float32)
    gg = np.ascontiguousarray(g.gamma.ravel(order="F")).view(np.float32)
    ulps = np.abs(gp.view(np.int32).astype(np.int64) - gg.view(np.int32).astype(np.int64))
    ulps[(gp == 0) & (gg == 0)] = 0  # +0 and -0
    assert ulps.max() <= 1 and np.count_nonzero(ulps) <= max(1, ulps.size // 10000), (ulps.max(), np.count_nonzero(ulps))
    assert len(p["wiggles"]) == len(g.wiggles)
    for a, b in zip(p["wiggles"], g.wiggles):
        for k in ("pxe", "pye", "pze", "gx", "gy", "gz"):
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("builder", [cases.case_2d_homogeneous, cases.case_2d_disordered, cases.case_3d_disordered,
                                     cases.case_1d, cases.case_periodic_2d])
def test_scalar_setup_equals_the_python_mirror(builder):
    n, opt, _ = builder()
    _compare(n, opt)


def test_vector_setup_equals_the_python_mirror():
    n, opt, _ = cases.case_3d_disordered()
    _compare(n, opt, vector=True)


@pytest.mark.parametrize("shape,bw", [((25, 49), [7, 5]), ((13, 11, 9), [4, 6, 5]), ((40,), [9, 0]),
                                      ((16, 20, 12), [0, 6, 4])])
def test_odd_sizes_and_mixed_boundaries(shape, bw):
    rng = np.random.default_rng(7)
    n = 1.0 + 0.4 * rng.random(shape) + 0.02j * rng.random(shape)
    _compare(n, dict(pixel_size=0.25, boundary_widths=bw))


def test_options_epsilon_roi_acc_lambda():
    rng = np.random.default_rng(8)
    n = 1.2 + 0.3 * rng.random((20, 18))
    _compare(n, dict(pixel_size=0.2, boundary_widths=[6, 6], epsilon=12.5))
    _compare(n, dict(pixel_size=0.2, boundary_widths=[6, 6], roi=[[2, 3], [15, 17]]))
    _compare(n, dict(pixel_size=0.2, boundary_widths=[6, 6], ACC=False))
    _compare(n, dict(pixel_size=0.2, boundary_widths=[6, 6], ACC=[True, False, False]))
    _compare(n, dict(pixel_size=0.1, boundary_widths=[6, 6], **{"lambda": 0.532}))


def test_errors_are_codes():
    from wavesim_amd import capi, problem
    with pytest.raises(capi.WsimError):
        problem.build(np.ones((4, 4)) - 0.1j, dict(pixel_size=0.25, boundary_widths=[2, 2]))
    with pytest.raises(capi.WsimError):
        problem.build(np.ones((4, 4)), dict(pixel_size=0.25, boundary_widths=[2, 2], roi=[[1, 1], [9, 9]]))
