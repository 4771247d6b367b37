"""Accuracy of the vector solver against Mie theory (run on a GPU host).

    python -m tests.mie_report reference     # the reference's Mie example, boundaries 2..8 wavelengths
    python -m tests.mie_report reference_fine reference_finer   # the same sphere on 2x, 3x finer grids
    python -m tests.mie_report cfg3b         # the cfg3b sphere (n 1.6 in 1.33, radius 20 px, 128^3 + 32 px)

Each case restates examples/mie_sphere/mie_sphere_test.m with the GPU solver
(tests/mie_cases.py) and the Mie oracle (oracle/mie.py) and prints one JSON
line: the reference's error measure mean|E_sim - E_theory|^2 / mean|E_theory|^2
(mie_sphere_test.m:49), its square root (a relative L2 error) and the
iteration count.  This is a report, not a test: the oracle evaluates the Mie
series on the host.
"""
from __future__ import annotations

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

CASES = {
    # mie_sphere_analytical_solution.m:19-26: lambda 1 um, 5 points per wavelength, radius 1 um,
    # n 1.2 in 1.0, grid [60 40 30]; boundary widths in wavelengths (mie_sphere_test.m:29)
    "reference": dict(N=[60, 40, 30], ppw=5, radius=1.0, n_sphere=1.2, n_bg=1.0, widths=[2, 4, 6, 8]),
    # the same with twice the iterations (convergence check)
    "reference_long": dict(N=[60, 40, 30], ppw=5, radius=1.0, n_sphere=1.2, n_bg=1.0, widths=[6], cycles=[400]),
    # the same geometry on grids 2x and 3x finer (discretisation check)
    "reference_fine": dict(N=[120, 80, 60], ppw=10, radius=1.0, n_sphere=1.2, n_bg=1.0, widths=[6]),
    "reference_finer": dict(N=[180, 120, 90], ppw=15, radius=1.0, n_sphere=1.2, n_bg=1.0, widths=[6]),
    # cfg3b: 128^3 interior, 0.25 um pixels (4 per wavelength), 32 px = 8 wavelength boundary,
    # sphere radius 20 px = 5 um, n 1.6 in 1.33
    "cfg3b": dict(N=[128, 128, 128], ppw=4, radius=5.0, n_sphere=1.6, n_bg=1.33, widths=[8]),
}


def run(name):
    from tests import mie_cases as mc
    c = CASES[name]
    t0 = time.perf_counter()
    Et = mc.theory(c["N"], c["ppw"], c["radius"], c["n_sphere"], c["n_bg"])
    t_theory = time.perf_counter() - t0
    for w in c["widths"]:
        for cycles in c.get("cycles", [200]):
            t1 = time.perf_counter()
            Es, st, (X, Y, Z), Ebg = mc.simulate(c["N"], c["ppw"], c["radius"], c["n_sphere"], c["n_bg"], w,
                                                 max_cycles=cycles)
            err = mc.error(Es, Et)
            inc = mc.incident_deviation(Ebg, X, Y, Z, c["radius"], c["n_bg"])
            print(json.dumps({"case": name, "grid": c["N"], "boundary_wavelengths": w, "max_cycles": cycles,
                              "iterations": st.iterations, "error_mean_sq": err, "rel_l2": err ** 0.5,
                              "incident_beam_vs_plane_wave_rel_l2": inc,
                              "gpu_s": round(time.perf_counter() - t1, 2), "theory_s": round(t_theory, 1)}),
                  flush=True)


def main(argv=None):
    for n in (argv if argv is not None else sys.argv[1:]) or ["reference"]:
        run(n)


if __name__ == "__main__":
    main()
