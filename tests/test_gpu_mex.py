"""The MEX shim RunWaveSimAlgorithmMex run end to end through a CPU fake of
MATLAB's C API (tests/mex_stub): the 12 mxArray inputs WaveSimBase hands it
(WaveSimBase.m:145-146), the HIP loop through the C ABI, and the outputs the
caller post-processes (WaveSimBase.m:150-155): E' on the full grid and
diff_energy of one entry per iteration, against the oracle; progress
forwarded to SimulationCallback(callbackType, diff_energy, energy,
energy_threshold) (SimulationCallback.m:1, 21-28) after every multiple of the
interval and at the end."""
from __future__ import annotations

import math

import numpy as np
import pytest

from tests import cases
from tests.cases import rel_l2




def _harness():
    from tests.mex_stub.harness import Mex
    return Mex()


def _run(builder, vector=False, **kw):
    from oracle import mbs
    n, opt, src = builder()
    o = mbs.make_sim(n, opt, vector=vector)
    s = mbs.prepare_source(o, cases.oracle_source(src))
    mx = _harness()
    prhs = mx.inputs(o, s, **kw)
    try:
        E, de = mx.call(prhs)
    finally:
        mx.free(prhs)
    return o, cases.oracle_source(src), mx, E, de


@pytest.mark.gpu
@pytest.mark.parametrize("builder,vector", [(cases.case_2d_disordered, False), (cases.case_3d_vector_sphere, True),
                                            (cases.case_3d_disordered, False)])
def test_mex_outputs_match_oracle(gpu, builder, vector):
    from oracle import mbs
    o, src, mx, E, de = _run(builder, vector)
    Eo, so = mbs.run(o, src)
    de = de.ravel()
    it = de.size + 1                              # WaveSimBase.m:154
    converged = it < 100000                       # WaveSimBase.m:155 (maxIter = 100000)
    print(f"{builder.__name__}: E' {E.shape}, {de.size} iterations (oracle {so.it - 1}), "
          f"rel-L2 {rel_l2(E.reshape(o.N), so.E_primed):.2e}", flush=True)
    assert E.dtype == np.complex64 and E.size == np.prod(o.N)
    assert de.size == so.it - 1 and converged == so.converged
    assert rel_l2(E.reshape(o.N), so.E_primed) <= 1e-5
    assert np.allclose(de, so.diff_energy, rtol=1e-4)
    with np.errstate(all="ignore"):                # the caller's E = crop(E' ./ gamma), WaveSimBase.m:150-153
        Ec = mbs.crop(o, E.reshape(o.N) / o.gamma[..., None])
    assert rel_l2(Ec, Eo) <= 1e-5
    assert mx.L.mxf_live_temporaries() == 0


@pytest.mark.gpu
def test_mex_forwards_progress_to_simulation_callback(gpu):
    from oracle import mbs
    _harness().L.mxf_log_clear()
    o, _, mx, E, de = _run(cases.case_2d_disordered, callback="EnergyAddedDisp", interval=7.0)
    log = mx.log()
    n = de.size
    want = list(range(7, n + 1, 7))
    if want[-1] != n:
        want.append(n)
    print(f"{n} iterations, {len(log)} SimulationCallback calls", flush=True)
    assert len(log) == math.ceil(n / 7)
    assert [len(c["de"]) for c in log] == want
    for c in log:
        assert c["name"] == "SimulationCallback" and c["type"] == "EnergyAddedDisp" and c["nargs"] == 4
        assert c["energy_numel"] == 0 and c["threshold"] == o.energy_threshold
        assert np.array_equal(c["de"], de.ravel()[:len(c["de"])])


@pytest.mark.gpu
def test_mex_iteration_cap(gpu):
    """maxIter = 21 with Nw = 4: the cap acts at the next multiple of Nw
    (Simulation.m:202), so 24 entries, and the caller reports not converged."""
    _, _, _, E, de = _run(lambda: cases.case_2d_homogeneous(thr=1e-30), max_iter=21.0)
    assert de.size == 24 and not (de.size + 1 < 21)


@pytest.mark.skipif(__import__("tests.conftest", fromlist=["gpu_available"]).gpu_available(),
                    reason="checks the no-GPU path")
@pytest.mark.parametrize("vector", [False, True])
def test_mex_inputs_pass_the_shim_checks_without_gpu(vector):
    """On a host without a GPU the 12 inputs built like WaveSimBase's pass every
    argument check of the shim and reach the library, whose error (no device)
    comes back as a MATLAB error after the call, not as a crash."""
    from oracle import mbs
    from tests.mex_stub import build_shim
    import os
    if not os.path.exists(os.path.join(os.path.dirname(build_shim.__file__), "_build", "libmexshim.so")):
        build_shim.build()
    builder = cases.case_3d_vector_sphere if vector else cases.case_2d_disordered
    n, opt, src = builder()
    o = mbs.make_sim(n, opt, vector=vector)
    s = mbs.prepare_source(o, cases.oracle_source(src))
    mx = _harness()
    prhs = mx.inputs(o, s)
    try:
        with pytest.raises(RuntimeError, match="wsim:run"):
            mx.call(prhs)
    finally:
        mx.free(prhs)
    assert mx.L.mxf_live_temporaries() == 0
