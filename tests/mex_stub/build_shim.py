"""Builds the MEX shim against the CPU fake of MATLAB's C API (test
infrastructure).

``build()`` links ``wavesim_amd/mex/RunWaveSimAlgorithmMex.cpp`` with
``mxfake.cpp`` into ``tests/mex_stub/_build/libmexshim.so``, which loads
``wavesim_amd/libwavesim_mbs.so`` next to it in the tree: the GPU test calls
``mexFunction`` through it.  ``build_sanitized(out)`` compiles the host-only
C++ of the library -- the problem set-up ``csrc/problem.hip`` -- and the shim
with the fake, an in-process stand-in for ``wsim_run_algorithm`` and a
driver (``host_checks.cpp``) into one executable built with AddressSanitizer
and UndefinedBehaviorSanitizer, for the CPU suite.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
SHIM = os.path.join(ROOT, "wavesim_amd", "mex", "RunWaveSimAlgorithmMex.cpp")
FAKE = os.path.join(HERE, "mxfake.cpp")
LIBDIR = os.path.join(ROOT, "wavesim_amd")


def gxx() -> str:
    g = shutil.which("g++")
    if g is None:
        raise RuntimeError("g++ not found")
    return g


def build(out_dir: str = OUT) -> str:
    os.makedirs(out_dir, exist_ok=True)
    lib = os.path.join(out_dir, "libmexshim.so")
    rel = os.path.relpath(LIBDIR, out_dir)
    cmd = [gxx(), "-std=c++17", "-O2", "-Wall", "-fPIC", "-shared", "-I", HERE, "-I", os.path.join(ROOT, "include"),
           FAKE, SHIM, "-L", LIBDIR, "-lwavesim_mbs", f"-Wl,-rpath,$ORIGIN/{rel}", "-o", lib + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("MEX shim build failed:\n" + " ".join(cmd) + "\n" + r.stderr)
    os.replace(lib + ".tmp", lib)
    return lib


PROBLEM_SRC = os.path.join(ROOT, "wavesim_amd", "csrc", "problem" + ".hip")  # host-only C++
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


def build_sanitized(out_dir: str) -> str:
    exe = os.path.join(out_dir, "host_checks")
    # the set-up code is plain host C++: compiled by g++ from a .cpp copy
    host_cpp = os.path.join(out_dir, "problem_host.cpp")
    shutil.copyfile(PROBLEM_SRC, host_cpp)
    cmd = [gxx(), "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-Wno-unknown-pragmas"]
    cmd += SANITIZE
    cmd += ["-I", HERE, "-I", os.path.join(ROOT, "include"), host_cpp, FAKE, SHIM,
            os.path.join(HERE, "host_checks.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("sanitized host build failed:\n" + " ".join(cmd) + "\n" + r.stderr)
    return exe


if __name__ == "__main__":
    print(build())
