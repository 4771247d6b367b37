"""Groups of independent problems (ragged grids in shared launches) against
single-problem solves, and the first group member against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from tests import cases
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def _problems():
    """Ragged 2-D grids (several sizes, media and sources), a 1-D line and a
    3-D problem: different kernel sequences land in different sub-launches."""
    import wavesim_amd as w
    specs = [cases.case_2d_homogeneous(48, 8), cases.case_2d_disordered(40, 56, 8, 1),
             cases.case_2d_disordered(64, 30, 6, 5), cases.case_2d_homogeneous(96, 16),
             cases.case_2d_disordered(24, 24, 4, 9), cases.case_1d(64, 16),
             cases.case_3d_disordered((20, 24, 16), 4, 2)]
    out = []
    for n, opt, src in specs:
        out.append((w.WaveSim(n, opt), cases.gpu_source(src)))
    return specs, out


def test_group_equals_single_solves(gpu):
    import wavesim_amd as w
    specs, probs = _problems()
    grouped = w.exec_many(probs)
    for (sim, src), (Eg, sg) in zip(probs, grouped):
        Es, ss = sim.exec(src)
        assert sg.iterations == ss.iterations
        assert sg.converged == ss.converged
        # the same kernels on the same arguments: bit-identical
        assert np.array_equal(Eg, Es, equal_nan=True)
        assert np.array_equal(sg.diff_energy, ss.diff_energy)


def test_group_member_matches_oracle(gpu):
    import wavesim_amd as w
    from oracle import mbs
    specs, probs = _problems()
    grouped = w.exec_many(probs)
    for k in (1, 6):
        n, opt, src = specs[k]
        o = mbs.make_sim(n, opt)
        Eo, so = mbs.run(o, cases.oracle_source(src))
        assert grouped[k][1].iterations == so.it - 1
        assert rel_l2(grouped[k][0], Eo) <= 1e-5


def test_group_per_problem_limits(gpu):
    """max_iter per problem: a capped member stops at its cap (a multiple of
    Nw) while the others run to convergence."""
    import wavesim_amd as w
    specs, probs = _problems()
    svs = []
    for sim, src in probs[:3]:
        p = sim._prepare(src)
        sv = sim.solver(1)
        sv.set_source(0, p.positions, p.values)
        svs.append(sv)
    caps = [8, 100000, 100000]
    with w.SolverGroup(svs) as g:
        g.solve(caps, [s.energy_threshold for s, _ in probs[:3]])
    r0 = svs[0].result(0)
    assert r0.iterations == 8 and not r0.converged
    for (sim, src), sv in zip(probs[1:3], svs[1:]):
        r = sv.result(0)
        _, ss = sim.exec(src)
        assert r.iterations == ss.iterations and r.converged


def test_group_rejects_duplicates(gpu):
    import wavesim_amd as w
    from wavesim_amd.capi import WsimError
    _, probs = _problems()
    sv = probs[0][0].solver(1)
    with pytest.raises(WsimError):
        w.SolverGroup([sv, sv])


def test_group_with_warm_start_equals_single(gpu):
    """A warm-started member (the start is staged in the first pass before the
    grouped iterations) gives its single solve's result."""
    import wavesim_amd as w
    specs, probs = _problems()
    (sim0, src0), (sim1, src1) = probs[1], probs[3]
    rng = np.random.default_rng(4)
    E0 = (0.5 * rng.standard_normal(sim0.N) + 0.5j * rng.standard_normal(sim0.N)).astype(np.complex64)
    svs = []
    for sim, src in ((sim0, src0), (sim1, src1)):
        p = sim._prepare(src)
        sv = sim.solver(1)
        sv.set_source(0, p.positions, p.values)
        svs.append(sv)
    svs[0].set_warm_start(0, E0)
    svs[0].solve(sim0.max_iterations(), sim0.energy_threshold)
    single = svs[0].field(0).copy()
    it_single = svs[0].result(0).iterations
    with w.SolverGroup(svs) as g:
        g.solve([s.max_iterations() for s in (sim0, sim1)], [s.energy_threshold for s in (sim0, sim1)])
    assert svs[0].result(0).iterations == it_single
    assert np.array_equal(svs[0].field(0), single)
    svs[0].clear_warm_start()


def test_group_refuses_to_run_after_a_member_was_closed(gpu):
    import wavesim_amd as w
    _, probs = _problems()
    a, b = probs[0][0].solver(1), probs[1][0].solver(1)
    g = w.SolverGroup([a, b])
    b.close()
    with pytest.raises(RuntimeError, match="closed"):
        g.solve(8, 1e-6)
    g.close()
    with pytest.raises(RuntimeError, match="group is closed"):
        g.solve(8, 1e-6)
