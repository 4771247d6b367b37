"""How many MBS iterations a warm start saves, by the kind of error in its
initial field, on the GPU solver over several held-out media (report script;
tests/warmstart_sensitivity_report.py is the same study on the CPU oracle for
one medium).

Per medium of the learned warm start's workload (1-8 random spheres, 128^3
ROI, 192^3 grid, Nw = 8, plane wave): a tight solve (energy ratio 1e-14) gives
E*; the solver then restarts from E* spoiled one way at a time -- white or
smoothed relative noise, a global scale or phase error, E* with its Fourier
content above f k0 removed -- and from the straight-ray prior and the trained
network's E0, each to the zero start's absolute residual (fixed stop
denominator).  Prints the mean iterations and the ratio to the zero start.

    python -m tests.warmstart_sensitivity_gpu [n_media]
"""
from __future__ import annotations

import math
import sys

import numpy as np
import scipy.ndimage as ndi


def main(argv=None):
    import torch
    from wavesim_amd import WaveSim
    from wavesim_amd.learned import kernels as K
    from wavesim_amd.learned.warmstart import WarmStart
    from wavesim_amd.workloads import spheres_3d
    argv = sys.argv[1:] if argv is None else argv
    n_media = int(argv[0]) if argv else 8
    ws = WarmStart.load(__import__("wavesim_amd.learned.warmstart", fromlist=["x"]).DEFAULT_WEIGHTS)
    rows = {}
    zero_its = []
    for seed in range(n_media):
        nmap, opt, src = spheres_3d(128, 32, seed=seed)
        sim = WaveSim(nmap, opt)
        sv = sim.solver(1)
        p = sim._prepare(src)
        sv.set_source(0, p.positions, p.values)
        mi = sim.max_iterations()
        sv.solve(mi, sim.energy_threshold)
        r0 = sv.result(0)
        d1 = float(r0.diff_energy[0])
        zero_its.append(r0.iterations)
        sv.solve(mi, 1e-14)
        g = sim.gamma
        with np.errstate(all="ignore"):
            Et = np.where(g != 0, sv.field(0)[..., 0] / g, 0).astype(np.complex128)
        rng = np.random.default_rng(1 + seed)

        def noise():
            return rng.standard_normal(Et.shape) + 1j * rng.standard_normal(Et.shape)

        def smooth():
            nz = ndi.gaussian_filter(noise().real, 4) + 1j * ndi.gaussian_filter(noise().real, 4)
            return nz / np.sqrt(np.mean(np.abs(nz) ** 2))

        k0 = 2 * math.pi * 1.33
        F = np.fft.fftn(Et)
        pp = [np.fft.fftfreq(n, sim.grid.dx) * 2 * math.pi for n in Et.shape]
        P2 = pp[0][:, None, None] ** 2 + pp[1][None, :, None] ** 2 + pp[2][None, None, :] ** 2
        starts = {
            "exact": Et,
            "white noise 0.01": Et * (1 + 0.01 * noise() / math.sqrt(2)),
            "white noise 0.03": Et * (1 + 0.03 * noise() / math.sqrt(2)),
            "white noise 0.1": Et * (1 + 0.1 * noise() / math.sqrt(2)),
            "smooth noise 0.03": Et * (1 + 0.03 * smooth()),
            "smooth noise 0.1": Et * (1 + 0.1 * smooth()),
            "scale 0.97": Et * 0.97,
            "phase 0.05 rad": Et * np.exp(0.05j),
            "low-pass 1.5 k0": np.fft.ifftn(np.where(P2 <= (1.5 * k0) ** 2, F, 0)),
        }
        # the straight-ray prior and the trained network's prediction
        prep = ws.prepare(sim, nmap)
        zero = torch.zeros(2 * ws.cfg.L ** 3, dtype=torch.float32, device=ws.dev)
        for name, E0d in (("straight-ray prior", ws.compose(sim, prep, zero)),
                          ("learned (FNO)", ws.predict_E0(sim, prep))):
            torch.cuda.synchronize()
            starts[name] = E0d.cpu().numpy().view(np.complex64).reshape(sim.N[2], sim.N[1], sim.N[0]).transpose(
                2, 1, 0).astype(np.complex128)
        roi = tuple(slice(int(sim.roi[0, d]) - 1, int(sim.roi[1, d])) for d in range(3))
        for name, E0 in starts.items():
            sv.set_warm_start(0, E0[..., None])
            sv.set_reference_energy(0, d1)
            sv.solve(mi, sim.energy_threshold)
            its = sv.result(0).iterations
            err = float(np.linalg.norm((E0 - Et)[roi]) / np.linalg.norm(Et[roi]))
            rows.setdefault(name, []).append((its, err))
            sv.clear_warm_start()
            sv.set_reference_energy(0, 0.0)
        print(f"# medium {seed}: zero start {r0.iterations} iterations", flush=True)
    z = float(np.mean(zero_its))
    print(f"{'start':22s} {'error':>8s} {'iterations':>11s} {'zero/start':>11s}   (mean over {n_media} media; "
          f"zero start {z:.1f} iterations)")
    for name, v in rows.items():
        its = np.mean([a for a, _ in v])
        err = np.mean([b for _, b in v])
        print(f"{name:22s} {err:8.4f} {its:11.1f} {z / its:11.2f}")


if __name__ == "__main__":
    main()
