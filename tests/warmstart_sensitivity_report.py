"""How many MBS iterations a warm start saves, by the kind of error in its
initial field (report script on the CPU oracle; no GPU needed).

For one held-out medium of the learned warm start's workload (1-8 random
spheres, 128^3 ROI, 192^3 grid, Nw = 8, plane wave), the oracle solves to a
tight residual (energy ratio 1e-14) and then restarts from that solution E*
spoiled in one way at a time: E* times (1 + a noise) with white or smoothed
noise, a global scale or phase error, and E* with its Fourier content above
f k0 removed.  Each restart runs to the zero start's absolute residual (the
zero start's first increment energy as the fixed denominator) and reports the
iterations it needs.  The warm start enters through the increment recurrence
as in the library (oracle.mbs.run with E0).

    python -m tests.warmstart_sensitivity_report [seed]
"""
from __future__ import annotations

import math
import os
import sys
import time

import numpy as np
import scipy.ndimage as ndi


def main(argv=None):
    from oracle import mbs
    from wavesim_amd.workloads import spheres_3d
    argv = sys.argv[1:] if argv is None else argv
    seed = int(argv[0]) if argv else 0
    workers = os.cpu_count()
    nmap, opt, src = spheres_3d(128, 32, seed=seed)
    sim = mbs.make_sim(nmap, opt)
    S = mbs.Source(list(src.positions), [np.asarray(v) for v in src.values])
    t0 = time.perf_counter()
    _, sz = mbs.run(sim, S, workers=workers)
    d1 = sz.diff_energy[0]
    print(f"seed {seed}: zero start {sz.it - 1} iterations ({time.perf_counter() - t0:.0f} s)", flush=True)
    tight = mbs.make_sim(nmap, dict(opt, energy_threshold=1e-14))
    _, st = mbs.run(tight, S, workers=workers)
    g = sim.gamma
    with np.errstate(all="ignore"):
        E0t = np.where(g != 0, st.E_primed[..., 0] / g, 0)
    roi = tuple(slice(int(sim.roi[0, d]) - 1, int(sim.roi[1, d])) for d in range(3))
    rng = np.random.default_rng(1)

    def noise(shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    def run(name, E0):
        t = time.perf_counter()
        _, s = mbs.run(sim, S, workers=workers, E0=E0[..., None], ref_energy=d1)
        err = float(np.linalg.norm((E0 - E0t)[roi]) / np.linalg.norm(E0t[roi]))
        print(f"{name:22s} start error {err:.4f}  iterations {s.it - 1:4d}  "
              f"(zero start {sz.it - 1}, ratio {(sz.it - 1) / (s.it - 1):.2f}; {time.perf_counter() - t:.0f} s)",
              flush=True)

    run("exact", E0t)
    for a in (0.01, 0.03, 0.1):
        run(f"white noise {a}", E0t * (1 + a * noise(E0t.shape) / math.sqrt(2)))
    for a in (0.03, 0.1):
        nz = ndi.gaussian_filter(noise(E0t.shape).real, 4) + 1j * ndi.gaussian_filter(noise(E0t.shape).real, 4)
        nz /= np.sqrt(np.mean(np.abs(nz) ** 2))
        run(f"smooth noise {a}", E0t * (1 + a * nz))
    run("scale 0.97", E0t * 0.97)
    run("phase 0.05 rad", E0t * np.exp(0.05j))
    k0 = 2 * np.pi * 1.33
    F = np.fft.fftn(E0t)
    p = [np.fft.fftfreq(n, 0.25) * 2 * np.pi for n in E0t.shape]
    P2 = p[0][:, None, None] ** 2 + p[1][None, :, None] ** 2 + p[2][None, None, :] ** 2
    for f in (1.5, 2.0):
        run(f"low-pass {f} k0", np.fft.ifftn(np.where(P2 <= (f * k0) ** 2, F, 0)))


if __name__ == "__main__":
    main()
