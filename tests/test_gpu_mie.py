"""The vector solver against Mie theory on the reference's Mie-sphere example
(examples/mie_sphere/mie_sphere_test.m with ARL boundaries and ACC, 6
wavelengths wide, the width of the reference's cross-section figure,
mie_sphere_test.m:57-66).

Bounds: the reference's error measure mean|E_sim - E_theory|^2 /
mean|E_theory|^2 was 0.0202 at 5 points per wavelength and 0.0047 at 10 when
measured on MI355X.  The finer grid's error is set by the example's tapered
source beam (5.5% from a plane wave near the sphere), not by the solver.
"""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

GEOMETRY = dict(radius=1.0, n_sphere=1.2, n_bg=1.0)


def test_mie_sphere_reference_example_and_refinement(gpu):
    from tests import mie_cases as mc
    errs = {}
    for N, ppw in (([60, 40, 30], 5), ([120, 80, 60], 10)):
        Et = mc.theory(N, ppw, **GEOMETRY, workers=1)
        Es, st, (X, Y, Z), Ebg = mc.simulate(N, ppw, GEOMETRY["radius"], GEOMETRY["n_sphere"], GEOMETRY["n_bg"], 6)
        errs[ppw] = mc.error(Es, Et)
        inc = mc.incident_deviation(Ebg, X, Y, Z, GEOMETRY["radius"], GEOMETRY["n_bg"])
        print(f"{ppw} points per wavelength: {st.iterations} iterations, error {errs[ppw]:.4f}, "
              f"incident beam vs plane wave {inc:.3f}", flush=True)
        assert st.iterations == 136  # 200 cycles, rounded up to a multiple of Nw = 8
    assert errs[5] <= 0.021
    assert errs[10] <= 0.005
