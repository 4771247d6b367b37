"""PSTD (PSTD.m) on the HIP axis passes against the oracle restatement
(oracle/pstd.py): same number of steps, same stop, the phase-compensated
field on the ROI to rel-L2 <= 1e-5 in complex64, step energies to 1e-4.
Media are run with boundary_type='ARL' as PSTD.m:32-34 requires (the medium is
replicate-padded to the FFT grid; absorption comes from Im n)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from tests import cases
from tests.cases import rel_l2

TOL = 1e-5


def _opt(bw, thr=1e-5, **kw):
    o = dict(pixel_size=0.25, boundary_widths=bw, boundary_type="ARL", energy_threshold=thr,
             callback_interval=50)
    o.update(kw)
    return o


def _disordered_2d():
    n = cases.smooth_disorder((40, 52), 11, absorption=0.03)
    return n, _opt([6, 6]), [(1.0, [20, 26])]


def _disordered_3d():
    n = cases.smooth_disorder((16, 20, 12), 12, absorption=0.04)
    return n, _opt([4, 4, 4], 1e-4), [(1.0, [8, 10, 6])]


def _plane_wave_2d():
    n = cases.sphere((36, 36), 8, 1.3 + 0.02j, 1.0 + 0.02j)
    return n, _opt([8, 8], 1e-4, **{"lambda": 0.8}), [(np.ones(36), [1, 1])]


def _line_1d():
    n = np.full((64, 1), 1.2 + 0.03j)
    return n, _opt([8, 0], 1e-6), [(1.0, [32, 1])]


def _sources(src):
    import wavesim_amd as w
    from oracle import mbs
    g, o = None, None
    for v, p in src:
        gs, os_ = w.Source(v, p), mbs.Source.point(v, p)
        g = gs if g is None else g + gs
        o = os_ if o is None else o + os_
    return g, o


def test_host_setup_matches_oracle():
    """The package's PSTD constructor computes the oracle's grid, ROI, dt and
    coefficient maps (no GPU needed)."""
    import wavesim_amd as w
    from oracle import pstd
    for builder in (_disordered_2d, _disordered_3d, _plane_wave_2d, _line_1d):
        n, opt, _ = builder()
        g, o = w.PSTD(n, opt), pstd.make_pstd(n, opt)
        assert g.N == o.N and np.array_equal(g.roi, o.roi)
        assert g.dt == o.dt and g.iterations_per_cycle == o.iterations_per_cycle
        for a, b in ((g.c1, o.c1), (g.c2, o.c2), (g.c3, o.c3)):
            assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        w.PSTD(np.ones((8, 8)), dict(pixel_size=0.25, boundary_widths=[2, 2]))


@pytest.mark.gpu
@pytest.mark.parametrize("builder", [_disordered_2d, _disordered_3d, _plane_wave_2d, _line_1d])
def test_pstd_matches_oracle(gpu, builder):
    import wavesim_amd as w
    from oracle import pstd
    n, opt, src = builder()
    gs, os_ = _sources(src)
    o = pstd.make_pstd(n, opt)
    Eo, so = pstd.run(o, os_)
    g = w.PSTD(n, opt)
    Eg, sg = g.exec(gs)
    print(f"{builder.__name__}: steps oracle={so.it - 1} gpu={sg.iterations}, "
          f"converged {so.converged}/{sg.converged}, rel-L2 {rel_l2(Eg, Eo):.2e}", flush=True)
    assert sg.iterations == so.it - 1 and sg.converged == so.converged
    assert rel_l2(Eg, Eo) <= TOL
    assert np.allclose(sg.diff_energy, so.diff_energy, rtol=1e-4)
    g.close()


@pytest.mark.gpu
def test_pstd_periodic_steady_state_closed_form(gpu):
    """Homogeneous absorbing periodic medium, 3000 steps: the GPU field equals
    the recurrence's time-harmonic closed form X / z (p = 0 offset excluded)."""
    import wavesim_amd as w
    from oracle import mbs, pstd
    n = np.full((24, 32), 1.0 + 0.05j)
    opt = _opt([0, 0], 0.0)
    g = w.PSTD(n, opt)
    E, st = g.exec(w.Source(1.0, [12, 16]), max_iterations=3000)
    assert not st.converged and st.iterations >= 3000
    o = pstd.make_pstd(n, dict(opt, single_precision=False))
    X = pstd.steady_state_periodic(o, mbs.Source.point(1.0, [12, 16]).dense(o.N, np.complex128))
    a, b = E - E.mean(), X - X.mean()
    err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"periodic steady state: {st.iterations} steps, rel-L2 vs closed form {err:.2e}", flush=True)
    assert err <= 2e-5
    g.close()


@pytest.mark.gpu
def test_pstd_iteration_cap_and_reuse(gpu):
    """max_iterations applies only on steps that may terminate (real A > 0.9),
    as Simulation.next with PSTD's can_terminate; a second exec on the same
    handle starts again from E = E_prev = 0 and equals the first."""
    import wavesim_amd as w
    from oracle import pstd
    n, opt, src = _disordered_2d()
    gs, os_ = _sources(src)
    o = pstd.make_pstd(n, dict(opt, energy_threshold=0.0))
    Eo, so = pstd.run(o, os_, max_iterations=900)
    g = w.PSTD(n, dict(opt, energy_threshold=0.0))
    E1, s1 = g.exec(gs, max_iterations=900)
    E2, s2 = g.exec(gs, max_iterations=900)
    assert s1.iterations == so.it - 1 and not s1.converged
    assert np.array_equal(E1, E2)
    assert rel_l2(E1, Eo) <= TOL
    g.close()


@pytest.mark.gpu
def test_pstd_on_a_shipped_grid_runs_specialised_kernels(gpu, monkeypatch):
    """A PSTD solver on a grid with a shipped MBS plan (the 128^3 spheres
    medium, 192^3 grid) runs that plan, its x and z passes and its fused
    update on plan-specialised kernels: the same steps and the field of the
    run-time-plan kernels on the same plan to 1e-5."""
    from wavesim_amd import PSTD
    from wavesim_amd.workloads import WORKLOADS
    nmap, opt, src = WORKLOADS["spheres_3d_128"]()
    popt = dict(opt, boundary_type="ARL", energy_threshold=0.0)
    out = []
    for mode in ("1", "0"):
        monkeypatch.setenv("WSIM_FIXED", mode)
        p = PSTD(np.asarray(nmap) + 0.02j, popt)
        nf = p.fixed_launches
        E, st = p.exec(src, max_iterations=120)
        out.append((nf, st.iterations, E))
        p.close()
    (nf1, it1, E1), (nf0, it0, E0) = out
    print(f"PSTD 192^3: fixed launches {nf1}, steps {it1} / {it0}, rel-L2 {rel_l2(E1, E0):.2e}", flush=True)
    assert nf1 == 4 and nf0 == 0
    assert it1 == it0 > 0
    assert rel_l2(E1, E0) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("builder", [_disordered_2d, _disordered_3d])
def test_pstd_fused_update_equals_separate_passes(gpu, builder, monkeypatch):
    """The update fused with the next step's first forward pass (the
    default) and the two as separate launches (WSIM_FUSE=0) compute the same
    steps: same step count and stop, fields equal to rounding."""
    import wavesim_amd as w
    n, opt, src = builder()
    gs, _ = _sources(src)
    out = []
    for mode in ("1", "0"):
        monkeypatch.setenv("WSIM_FUSE", mode)
        g = w.PSTD(n, opt)
        E, st = g.exec(gs)
        out.append((E, st))
        g.close()
    (E1, s1), (E0, s0) = out
    print(f"{builder.__name__}: steps {s1.iterations} / {s0.iterations}, rel-L2 {rel_l2(E1, E0):.2e}", flush=True)
    assert s1.iterations == s0.iterations and s1.converged == s0.converged
    assert rel_l2(E1, E0) <= 1e-6
