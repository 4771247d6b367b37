"""The committed golden fixtures (tests/golden/, written by the oracle with
tests/golden/make_fixtures.py): the oracle still reproduces them (CPU), and the
GPU solver matches them (GPU) without running the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from tests import cases
from tests.cases import rel_l2
from tests.golden.make_fixtures import SOLVES, load

NAMES = sorted(SOLVES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    from oracle import mbs
    nmap, opt, src, vector, d = load(name)
    o = mbs.make_sim(nmap, opt, vector=vector)
    E, st = mbs.run(o, cases.oracle_source(src))
    assert st.it - 1 == int(d["iterations"])
    assert rel_l2(E, d["E"]) <= 1e-6
    if "prop_in" in d:
        for w in d["prop_wiggles"]:
            assert rel_l2(mbs.propagate(o, d["prop_in"], o.wiggles[int(w)]), d[f"prop_out_{int(w)}"]) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(gpu, name):
    import wavesim_amd as w
    nmap, opt, src, vector, d = load(name)
    sim = (w.WaveSimVector if vector else w.WaveSim)(nmap, opt)
    E, st = sim.exec(cases.gpu_source(src))
    print(f"{name}: iterations gpu={st.iterations} fixture={int(d['iterations'])}", flush=True)
    assert st.iterations == int(d["iterations"])
    assert st.converged == bool(d["converged"])
    assert rel_l2(E, d["E"]) <= 1e-5
    assert np.allclose(st.diff_energy, d["diff_energy"], rtol=1e-4)
    if "prop_in" in d:
        for wg in d["prop_wiggles"]:
            assert rel_l2(sim.propagate(d["prop_in"], int(wg)), d[f"prop_out_{int(wg)}"]) <= 1e-5
