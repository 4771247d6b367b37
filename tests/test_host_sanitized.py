"""The library's host-only C++ (the problem set-up, csrc/problem.hip) and the
MEX shim's marshalling through the fake MATLAB API, built with
AddressSanitizer and UndefinedBehaviorSanitizer and run on the CPU
(tests/mex_stub/host_checks.cpp): every check passes and no sanitizer reports
an error or a leak."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    sys.path.insert(0, os.path.join(HERE, "mex_stub"))
    try:
        import build_shim
    finally:
        sys.path.pop(0)
    exe = build_shim.build_sanitized(str(tmp_path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host checks passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
