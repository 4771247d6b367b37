"""The allocation-free entry point wsim_run_algorithm_noalloc against
wsim_run_algorithm (the existing drop-in it varies), through the C ABI: the
same iterations, diff_energy and E' (rel-L2 <= 1e-6, complex64), on two calls
in a row -- the second on a smaller problem, so it runs on the scratch array
as the first call left it.  Device arrays for gamma and E' come from torch
(the caller's memory); the entry point itself allocates nothing."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests import cases
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def _inputs(n, opt, src):
    from oracle import mbs
    from wavesim_amd import capi
    o = mbs.make_sim(n, opt)
    s = mbs.prepare_source(o, cases.oracle_source(src))
    return o, capi.Wiggles(o.wiggles), capi.Patches(s.positions, s.values)


def _drop_in(o, W, P):
    from wavesim_amd import capi
    L = capi.lib()
    N = (C.c_double * 4)(*o.N)
    roi = (C.c_double * 8)(*np.asarray(o.roi, np.float64).T.ravel())
    g = capi.c64(o.gamma)
    E = np.zeros(o.N, np.complex64, order="F")
    de = (C.c_double * 100000)()
    nde = C.c_int64()
    capi.check(L.wsim_run_algorithm(N, capi.ptr(g), W.n, W.arr, P.n, P.arr, float(o.epsilon),
                                    capi.cfloat(o.k02e.real, o.k02e.imag), roi, float("inf"),
                                    o.energy_threshold, b"NoCallback", 50.0, capi.NULL_CALLBACK, None,
                                    capi.ptr(E), de, 100000, C.byref(nde)))
    return E, np.asarray(de[:nde.value])


def _no_alloc(o, W, P):
    import torch
    from wavesim_amd import capi
    L = capi.lib()
    N = (C.c_double * 4)(*o.N)
    roi = (C.c_double * 8)(*np.asarray(o.roi, np.float64).T.ravel())
    # gamma and E' in [y x z pol] column-major order = the reversed C order
    g = torch.from_numpy(np.ascontiguousarray(o.gamma.astype(np.complex64).T)).to("cuda")
    E = torch.full(tuple(reversed(o.N)), complex(7.0, -3.0), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    de = (C.c_double * 100000)()
    nde = C.c_int64()
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.wsim_run_algorithm_noalloc(N, g.data_ptr(), W.n, W.arr, P.n, P.arr, float(o.epsilon),
                                      capi.cfloat(o.k02e.real, o.k02e.imag), roi, float("inf"),
                                      o.energy_threshold, b"NoCallback", 50.0, capi.NULL_CALLBACK, None,
                                      E.data_ptr(), de, 100000, C.byref(nde), stream)
    capi.check(rc)
    return E.cpu().numpy().T, np.asarray(de[:nde.value])


def test_no_alloc_matches_drop_in_twice_second_smaller(gpu):
    problems = [cases.case_2d_disordered(ny=64, nx=80), cases.case_2d_disordered(ny=24, nx=32, seed=4)]
    for k, (n, opt, src) in enumerate(problems):
        o, W, P = _inputs(n, opt, src)
        E1, d1 = _drop_in(o, W, P)
        E2, d2 = _no_alloc(o, W, P)
        e = rel_l2(E2, E1)
        print(f"call {k}: grid {o.N}, iterations {len(d1)} / {len(d2)}, rel-L2 {e:.2e}, "
              f"bit-identical {np.array_equal(E1, E2)}", flush=True)
        assert len(d2) == len(d1) > 0
        assert e <= 1e-6
        assert np.allclose(d2, d1, rtol=1e-6)


def test_no_alloc_3d_after_2d(gpu):
    n, opt, src = cases.case_3d_disordered()
    o, W, P = _inputs(n, opt, src)
    E1, d1 = _drop_in(o, W, P)
    E2, d2 = _no_alloc(o, W, P)
    print(f"3-D: iterations {len(d1)} / {len(d2)}, rel-L2 {rel_l2(E2, E1):.2e}", flush=True)
    assert len(d2) == len(d1) and rel_l2(E2, E1) <= 1e-6


def test_no_alloc_refuses_oversized_problems(gpu):
    """Beyond the 64 MiB of work fields, and beyond 100 000 iterations."""
    import torch
    from wavesim_amd import capi
    L = capi.lib()
    n, opt, src = cases.case_2d_disordered()
    o, W, P = _inputs(n, opt, src)
    roi = (C.c_double * 8)(*np.asarray(o.roi, np.float64).T.ravel())
    g = torch.zeros(8, dtype=torch.complex64, device="cuda")
    de = (C.c_double * 8)()
    nde = C.c_int64()
    big = (C.c_double * 4)(1152, 1152, 4, 1)  # 3 x 8 B x 5.3 M voxels > 64 MiB
    rc = L.wsim_run_algorithm_noalloc(big, g.data_ptr(), W.n, W.arr, P.n, P.arr, float(o.epsilon),
                                      capi.cfloat(1.0, 1.0), roi, float("inf"), 1e-3, b"NoCallback", 50.0,
                                      capi.NULL_CALLBACK, None, g.data_ptr(), de, 8, C.byref(nde), None)
    assert rc == -3 and b"MiB" in L.wsim_last_error()
    N = (C.c_double * 4)(*o.N)
    rc = L.wsim_run_algorithm_noalloc(N, g.data_ptr(), W.n, W.arr, P.n, P.arr, float(o.epsilon),
                                      capi.cfloat(1.0, 1.0), roi, 2e5, 1e-3, b"NoCallback", 50.0,
                                      capi.NULL_CALLBACK, None, g.data_ptr(), de, 8, C.byref(nde), None)
    assert rc == -3 and b"iterations" in L.wsim_last_error()
