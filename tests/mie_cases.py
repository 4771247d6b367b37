"""The reference's Mie-sphere comparison restated for the GPU solver.

examples/mie_sphere/run_mie_sphere_simulation.m:3-31 simulates a sphere
(create_sphere.m) lit by an x-polarised plane wave entering at the first z
plane, subtracts a second simulation without the sphere (same epsilon), and
mie_sphere_test.m:49 compares the difference with the Mie scattered field of
mie_sphere_analytical_solution.m (here oracle/mie.py) by
mean(|E_sim - E_theory|^2) / mean(|E_theory|^2).  The steps are kept as the
reference writes them, including that the source window's x profile has
size(n, 1) samples and its y profile size(n, 2) (run_mie_sphere_simulation.m:
14-17; the same thing on a square cross-section).
"""
from __future__ import annotations

import numpy as np
from scipy.signal import windows


def simulate(N, ppw, radius, n_sphere, n_bg, boundary_wavelengths, lam=1.0, max_cycles=200):
    """(E_scattered [y,x,z,pol] from the GPU solver, solver state, (X, Y, Z),
    E_background)."""
    import wavesim_amd as w
    from oracle import mie
    ps = lam / ppw
    nmap = mie.sphere_index(N, ps, radius, n_sphere, n_bg)
    # mie_sphere_test.m:21-26, 41
    opt = {"pixel_size": ps, "lambda": lam, "boundary_widths": [boundary_wavelengths * lam / ps] * 3,
           "energy_threshold": 1e-99, "max_cycles": max_cycles, "callback_interval": 8}
    sim = w.WaveSimVector(nmap, opt)
    k = n_bg * 2 * np.pi / lam
    prefactor = 1j * ps / (2 * k)
    _, _, _, (xr, yr, zr) = mie.reference_grid(N, ps)
    sx, sy = nmap.shape[0], nmap.shape[1]
    srcx = windows.tukey(sx, 0.5).reshape(1, sx, 1)
    srcy = windows.tukey(sy, 0.5).reshape(sy, 1, 1)
    amp = (1 / prefactor) * np.exp(1j * k * zr[0]) * srcx * srcy
    src = w.Source(amp, [1, 1, 1, 1])
    E, st = sim.exec(src)
    bg = w.WaveSimVector(np.full(nmap.shape, n_bg), dict(opt, epsilon=float(sim.epsilon)))
    Ebg, _ = bg.exec(src)
    X, Y, Z, _ = mie.reference_grid(N, ps)
    return E - Ebg, st, (X, Y, Z), Ebg


def incident_deviation(Ebg, X, Y, Z, radius, n_bg, lam=1.0):
    """How far the simulated incident beam (the simulation without the sphere)
    is from the unit plane wave Mie theory assumes, within twice the sphere's
    radius (relative L2 over all components)."""
    k = 2 * np.pi * n_bg / lam
    near = X ** 2 + Y ** 2 + Z ** 2 < (2 * radius) ** 2
    plane = np.zeros(Ebg.shape, dtype=np.complex128)
    plane[..., 0] = np.exp(1j * k * Z)
    d = (Ebg - plane)[near]
    return float(np.linalg.norm(d) / np.linalg.norm(plane[near]))


def _slab(args):
    from oracle import mie
    radius, n_sphere, n_bg, lam, X, Y, Z = args
    return mie.near_field(radius, n_sphere, n_bg, lam, X, Y, Z)


def theory(N, ppw, radius, n_sphere, n_bg, lam=1.0, workers=None):
    """The Mie scattered field on the same grid ([y,x,z,pol]), evaluated in
    y slabs by a pool of host processes."""
    import concurrent.futures as cf
    import os
    from oracle import mie
    X, Y, Z, _ = mie.reference_grid(N, lam / ppw)
    workers = workers or min(16, os.cpu_count() or 1)
    slabs = np.array_split(np.arange(X.shape[0]), max(1, min(workers * 4, X.shape[0])))
    args = [(radius, n_sphere, n_bg, lam, X[s], Y[s], Z[s]) for s in slabs if s.size]
    if workers == 1:  # in-process (e.g. after the GPU runtime has started in this process)
        return np.concatenate([_slab(a) for a in args], axis=0)
    with cf.ProcessPoolExecutor(workers) as ex:
        parts = list(ex.map(_slab, args))
    return np.concatenate(parts, axis=0)


def error(E_sim, E_theory):
    """mie_sphere_test.m:49."""
    return float(np.mean(np.abs(E_sim - E_theory) ** 2) / np.mean(np.abs(E_theory) ** 2))
