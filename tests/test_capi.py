"""The C-ABI library: loads without a GPU, exports every symbol the header
declares, passes its host self-checks, and the package's host-side set-up
equals the oracle's (gamma, epsilon, k02e, wiggle tables)."""
import os
import re

import numpy as np
import pytest

from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|void)\s+(wsim_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared_symbols()
    assert "wsim_run_algorithm" in names and "wsim_solver_solve" in names
    assert len(names) >= 12


def test_library_exports_every_declared_symbol():
    from wavesim_amd import capi
    L = capi.lib()
    missing = [n for n in _declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.SYMBOLS) >= set(_declared_symbols())


def test_mex_shim_compiles_against_the_header():
    """The MEX shim (built on a MATLAB host) against include/wavesim_mbs.h and
    declarations of the MATLAB C API it calls: a syntax check, so that a change
    of the C ABI cannot leave the shim behind unnoticed."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "tests", "mex_stub"),
                        "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "wavesim_amd", "mex", "RunWaveSimAlgorithmMex.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_selftest_and_version():
    from wavesim_amd import capi
    L = capi.lib()
    assert L.wsim_abi_version() == 1
    assert L.wsim_selftest() == 0, L.wsim_last_error()


def test_errors_are_codes_not_crashes():
    import ctypes as C
    from wavesim_amd import capi
    L = capi.lib()
    h = C.c_void_p()
    assert L.wsim_solver_create(C.byref(h), None, None, None) == -1
    assert b"null" in L.wsim_last_error()
    assert L.wsim_solver_solve(None, 10, 0.1, 0, b"x", 0.0, capi.NULL_CALLBACK, None) == -1


def test_grid_size_limit_is_checked_before_any_device_work():
    """The passes address one component's field with a 32-bit byte offset, so
    a grid of more than 2^29 voxels is refused (WSIM_ERR_UNSUPPORTED) by the
    host-side check in create(), before any HIP call.  The descriptors carry
    an ROI outside the grid, the next host-side check: a grid the size check
    accepts fails there instead (WSIM_ERR_ARG), so no call reaches a device."""
    import ctypes as C
    from wavesim_amd import capi
    L = capi.lib()
    dummy = (C.c_float * 2)()
    wig = (capi.wiggle_t * 1)()

    def create(ny, nx, nz, C_=1):
        d = capi.desc_t()
        d.N[:] = [ny, nx, nz, C_]
        d.nrhs, d.nwiggles, d.epsilon, d.energy_interval = 1, 1, 3.0, 8
        d.roi[0][:] = [1, 1, 1, 1]
        d.roi[1][:] = [ny + 1, nx, nz, C_]  # outside the grid
        h = C.c_void_p()
        rc = L.wsim_solver_create(C.byref(h), C.byref(d), C.cast(dummy, C.c_void_p), wig)
        assert rc != 0
        return rc, L.wsim_last_error()

    for comps in (1, 3):
        rc, msg = create(1024, 1024, 512, comps)  # 2^29 voxels: the largest accepted
        assert rc == -1 and b"roi" in msg, (rc, msg)
        rc, msg = create(1024, 1024, 513, comps)
        assert rc == -3 and b"voxels" in msg, (rc, msg)

    # complex128 solvers: 16-byte elements, at most 2^28 voxels
    dummy_d = (C.c_double * 2)()

    def create_f64(ny, nx, nz, C_=1):
        d = capi.desc_f64_t()
        d.N[:] = [ny, nx, nz, C_]
        d.nrhs, d.nwiggles, d.epsilon, d.energy_interval = 1, 1, 3.0, 8
        d.roi[0][:] = [1, 1, 1, 1]
        d.roi[1][:] = [ny + 1, nx, nz, C_]
        h = C.c_void_p()
        rc = L.wsim_solver_create_f64(C.byref(h), C.byref(d), C.cast(dummy_d, C.c_void_p), wig)
        assert rc != 0
        return rc, L.wsim_last_error()

    for comps in (1, 3):
        rc, msg = create_f64(1024, 512, 512, comps)  # 2^28 voxels
        assert rc == -1 and b"roi" in msg, (rc, msg)
        rc, msg = create_f64(1024, 512, 513, comps)
        assert rc == -3 and b"voxels" in msg, (rc, msg)


@pytest.mark.parametrize("builder,vector", [(cases.case_2d_disordered, False),
                                            (cases.case_3d_disordered, False),
                                            (cases.case_1d, False),
                                            (cases.case_3d_vector_sphere, True)])
def test_host_setup_matches_oracle(builder, vector):
    import wavesim_amd as w
    from oracle import mbs
    n, opt, src = builder()
    o = mbs.make_sim(n, opt, vector=vector)
    g = (w.WaveSimVector if vector else w.WaveSim)(n, opt)
    assert g.N == o.N and np.array_equal(g.roi, o.roi)
    assert np.array_equal(g.gamma, o.gamma)
    assert g.epsilon == o.epsilon and g.k02e == o.k02e
    assert len(g.wiggles) == len(o.wiggles)
    for a, b in zip(g.wiggles, o.wiggles):
        for k in ("pxe", "pye", "pze", "gx", "gy", "gz"):
            assert np.array_equal(a[k], b[k]), k
    ps = g._prepare(cases.gpu_source(src))
    os_ = mbs.prepare_source(o, cases.oracle_source(src))
    assert [list(map(int, p)) for p in ps.positions] == [list(map(int, p)) for p in os_.positions]
    for a, b in zip(ps.values, os_.values):
        assert np.array_equal(a.astype(np.complex64), b.astype(np.complex64))


def test_efficient_size_product_matches_oracle():
    import wavesim_amd as w
    from oracle import mbs
    for s in range(1, 1200):
        assert w.efficient_size(s) == mbs.efficient_size(s)


def test_shipped_plans_become_a_compiled_table(tmp_path):
    """Every tuned plan becomes one initialiser row of the table compiled into
    the library; malformed entries stop the build."""
    import json
    from wavesim_amd import build, plans
    table = json.load(open(plans.SHIPPED))
    rows = [l for l in build.shipped_plans_inc().splitlines() if l.startswith("{")]
    assert len(rows) == len(table) > 0
    for k, v in table.items():
        m = re.fullmatch(r"(\d+)x(\d+)x(\d+)x(\d+)\|rhs(\d+)\|nw(\d+)(\|f64)?", k)
        f64 = 1 if m.group(7) else 0
        assert "{" + ", ".join(m.groups()[:6]) + f', {f64}, "{v["plan"]}"}},' in rows
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps({"320x320|rhs1|nw4": {"plan": "y:5,8,8/2"}}))
    with pytest.raises(ValueError):
        build.shipped_plans_inc(str(bad))
    bad.write_text(json.dumps({"320x320x1x1|rhs1|nw4": {"plan": 'y:5,8,8/2"); x'}}))
    with pytest.raises(ValueError):
        build.shipped_plans_inc(str(bad))
