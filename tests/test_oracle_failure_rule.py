"""The failure exit of the stop rule (a library addition: the reference runs a
diverging series to its cap) never fires on converging solves of the oracle,
and fires on the oracle's diverging ones (epsilon forced to half the
convergence bound, WaveSimBase.m:63), well before the cap."""
import numpy as np
import pytest

from tests import cases
from tests.cases import divergence_stop


@pytest.mark.parametrize("builder,vector", [(cases.case_2d_disordered, False), (cases.case_2d_homogeneous, False),
                                            (cases.case_3d_disordered, False), (cases.case_1d, False),
                                            (cases.case_periodic_2d, False), (cases.case_3d_vector_sphere, True)])
def test_rule_silent_on_converging_and_fires_on_diverging_series(builder, vector):
    from oracle import mbs
    n, opt, src = builder()
    o = mbs.make_sim(n, opt, vector=vector)
    nw = len(o.wiggles)
    _, s = mbs.run(o, cases.oracle_source(src), record_every=True)
    assert s.converged and divergence_stop(s.energy_every, nw) is None
    if float(o.epsilon) <= 3.0:
        return  # epsilon at its floor (epsilonmin): halving it still converges here
    o2 = mbs.make_sim(n, dict(opt, epsilon=0.5 * float(o.epsilon)), vector=vector)
    with np.errstate(all="ignore"):
        _, s2 = mbs.run(o2, cases.oracle_source(src), record_every=True, max_iterations=300)
    k = divergence_stop(s2.energy_every, nw)
    assert not s2.converged and k is not None and k <= 80, k
