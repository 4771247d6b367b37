"""The 256^3 disordered benchmark workload at its full size (320^3 grid with
its boundary layers), through properties that hold at any size: the oracle
would need minutes per iteration here, so these checks pin the full-size
pipeline (its tuned plan, tile shapes and launch counts) against exact
algebra and against rocFFT instead.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    from wavesim_amd.workloads import disordered_3d
    return disordered_3d(256, 32)


@pytest.fixture(scope="module")
def sim(full):
    import wavesim_amd as w
    nmap, opt, _ = full
    return w.WaveSim(nmap, opt)


def test_full_size_scaling_is_exact(gpu, full, sim):
    """Every step of the iteration is linear in (S, dE', E'), and doubling is
    exact in floating point: E(2S) == 2 E(S) bit for bit, with identical
    stopping decisions (the energies scale by 4 exactly).  The cap of 22 runs
    on a fresh solver and takes effect at iteration 24, the next multiple of
    Nw = 8 (Simulation.m:202)."""
    import wavesim_amd as w
    _, _, src = full
    pos = [int(p) for p in src.positions[0]]
    ran = []
    for cap in (22, 24):
        E1, s1 = sim.exec(w.Source(1.0, pos), max_iterations=cap)
        E2, s2 = sim.exec(w.Source(2.0, pos), max_iterations=cap)
        print(f"cap {cap}: {s1.iterations} iterations, |E| = {np.linalg.norm(E1):.6e}", flush=True)
        assert s1.iterations == s2.iterations
        ran.append(s1.iterations)
        assert np.array_equal(E2, 2 * E1)
        assert np.array_equal(np.asarray(s2.diff_energy), 4 * np.asarray(s1.diff_energy))
    assert ran == [24, 24]


def test_full_size_superposition(gpu, full, sim):
    """E(S1 + S2) = E(S1) + E(S2) for a fixed number of iterations (rounding
    only)."""
    import wavesim_amd as w
    n = full[0].shape[0]
    a = w.Source(1.0, [n // 2, n // 2, n // 2])
    b = w.Source(0.5 - 0.25j, [n // 3, 2 * n // 3, n // 4])
    Ea, _ = sim.exec(a, max_iterations=24)
    Eb, _ = sim.exec(b, max_iterations=24)
    Eab, _ = sim.exec(a + b, max_iterations=24)
    e = rel_l2(Eab, Ea + Eb)
    print(f"superposition rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-5


def test_full_size_fused_matches_separate_passes(gpu, full, sim, monkeypatch):
    """The fused pipeline (last inverse pass + next first forward pass, E'
    every other iteration) against separate passes at the benchmark size."""
    import wavesim_amd as w
    nmap, opt, src = full
    E1, s1 = sim.exec(src, max_iterations=24)
    monkeypatch.setenv("WSIM_FUSE", "0")
    E0, s0 = w.WaveSim(nmap, opt).exec(src, max_iterations=24)
    e = rel_l2(E1, E0)
    print(f"fused vs separate rel-L2 {e:.2e}", flush=True)
    assert s1.iterations == s0.iterations == 24
    assert e <= 1e-6
    assert np.allclose(s1.diff_energy, s0.diff_energy, rtol=1e-5)


def test_full_size_propagate_matches_rocfft(gpu, full, sim):
    """One application of the wiggled propagator (ramps, the three axis
    passes with the Green's function in the middle one, ramps undone) against
    torch.fft (rocFFT) with the oracle's tables, at 320^3."""
    import torch
    from oracle import mbs
    nmap, opt, _ = full
    o = mbs.make_sim(nmap, opt)
    rng = np.random.default_rng(11)
    E = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)).astype(np.complex64)
    for w in (0, len(o.wiggles) - 1):
        wig = o.wiggles[w]
        t = lambda a, shape: torch.from_numpy(np.asarray(a).astype(np.complex64).reshape(shape)).cuda()
        ramp = t(wig["gy"], (-1, 1, 1, 1)) * t(wig["gx"], (1, -1, 1, 1)) * t(wig["gz"], (1, 1, -1, 1))
        den = ((wig["pxe"] ** 2).reshape(1, -1, 1, 1) + (wig["pye"] ** 2).reshape(-1, 1, 1, 1)
               + (wig["pze"] ** 2 - o.k02e).reshape(1, 1, -1, 1))
        x = torch.from_numpy(E).cuda() * ramp
        x = torch.fft.ifftn(torch.fft.fftn(x, dim=(0, 1, 2)) / t(den, den.shape), dim=(0, 1, 2))
        ref = (x * ramp.conj()).cpu().numpy()
        del x
        e = rel_l2(sim.propagate(E, w), ref)
        print(f"wiggle {w}: rel-L2 vs rocFFT {e:.2e}", flush=True)
        assert e <= 1e-5


def test_cfg2_1024_disordered_full_size_matches_oracle(gpu):
    """The 1024^2 disordered workload at its full size (1152^2 grid, plane-wave
    line source) on its shipped plan -- whose x pass hands each XCD one range
    of tiles -- against the oracle: same iteration count, rel-L2 <= 1e-5."""
    import wavesim_amd as w
    from oracle import mbs
    from wavesim_amd.workloads import WORKLOADS
    nmap, opt, src = WORKLOADS["disordered_2d_1024"]()
    sim = w.WaveSim(nmap, opt)
    assert "/xcd" in sim.solver(1).plan
    Eg, sg = sim.exec(src)
    s = mbs.Source()
    for p, v in zip(src.positions, src.values):
        s = s + mbs.Source.point(np.asarray(v), list(p))
    Eo, so = mbs.run(mbs.make_sim(nmap, opt), s)
    e = rel_l2(Eg, Eo)
    print(f"1152^2: {sg.iterations} iterations, rel-L2 vs oracle {e:.2e}", flush=True)
    assert sg.iterations == so.it - 1 == 116
    assert e <= 1e-5


def test_cfg5_spheres_full_size_first_48_iterations_match_oracle(gpu):
    """The 128^3 spheres workload at its full size (192^3 grid, plane wave) on
    its shipped plan -- XCD tile ranges on the x passes -- against the oracle
    over the first 48 iterations (six wiggle cycles; the full 176 take the
    oracle about a minute)."""
    import wavesim_amd as w
    from oracle import mbs
    from wavesim_amd.workloads import WORKLOADS
    nmap, opt, src = WORKLOADS["spheres_3d_128"]()
    sim = w.WaveSim(nmap, opt)
    assert "/xcd" in sim.solver(1).plan
    Eg, sg = sim.exec(src, max_iterations=48)
    s = mbs.Source()
    for p, v in zip(src.positions, src.values):
        s = s + mbs.Source.point(np.asarray(v), list(p))
    Eo, so = mbs.run(mbs.make_sim(nmap, opt), s, max_iterations=48)
    e = rel_l2(Eg, Eo)
    print(f"192^3 spheres: {sg.iterations} iterations, rel-L2 vs oracle {e:.2e}", flush=True)
    assert sg.iterations == so.it - 1 == 48
    assert e <= 1e-5
