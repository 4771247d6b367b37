"""Build the HIP solver library in-tree: ``wavesim_amd/libwavesim_mbs.so``.

``python -m wavesim_amd.build`` compiles, for gfx950 with hipcc, the host
runtime ``csrc/mbs.hip`` and one object per kernel variant of
``csrc/variant.hip`` (in parallel), then links them.  Packed-FP32
instructions are disabled on the device side
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).  The FFT plans tuned
on MI355X (``tuned_plans.json``) are compiled into the library as a table
(``_build/shipped_plans.inc``), so every entry point -- the MEX-shaped
``wsim_run_algorithm`` included -- starts from them, and the pass launches of
those plans get kernels compiled with the plan as compile-time constants
(``fixed_plans.py``: one generated unit per launch shape plus a lookup table).
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwavesim_mbs.so")
FNO_LIB = os.path.join(HERE, "libwavesim_fno.so")
ARCH = os.environ.get("WSIM_OFFLOAD_ARCH", "gfx950")

# (contig, mode, prologue, epilogue) -- must match WSIM_FOR_EACH_VARIANT
VARIANTS = [(1, 0, 1, 0), (1, 1, 0, 1), (1, 1, 0, 3), (1, 1, 0, 2), (1, 2, 1, 1), (1, 2, 1, 2),
            (1, 1, 0, 4), (1, 1, 0, 5), (1, 2, 1, 4), (0, 0, 0, 0), (0, 1, 0, 0), (0, 2, 0, 0)]
COMPONENTS = (1, 3)
# kernel families: 0 complex64, 1 complex64 compact, 2 complex128 (the MBS
# passes only -- must match WSIM_FOR_EACH_VARIANT_F64)
FAMILY_F64 = 2
PSTD_EPILOGUES = (4, 5)  # the PSTD update, and the update fused with the next step's first pass


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def _flags():
    return ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC] + os.environ.get("WSIM_HIPCC_FLAGS", "").split()


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + msg)
    return msg


def shipped_plans_inc(path: str = os.path.join(HERE, "tuned_plans.json")) -> str:
    """C++ initialiser rows {Ny, Nx, Nz, C, nrhs, Nw, f64, "plan"} of the plans
    in ``path`` (keys as ``plans.key``: "NyxNxxNzxC|rhsR|nwW", "|f64" appended
    for complex128 solves)."""
    rows = ["// generated by wavesim_amd/build.py from tuned_plans.json"]
    with open(path) as f:
        table = json.load(f)
    for k in sorted(table):
        m = re.fullmatch(r"(\d+)x(\d+)x(\d+)x(\d+)\|rhs(\d+)\|nw(\d+)(\|f64)?", k)
        plan = table[k]["plan"]
        if not m or not re.fullmatch(r"[xyz0-9:,/;a-z]+", plan):
            raise ValueError(f"bad tuned plan entry {k!r}: {plan!r}")
        rows.append("{" + ", ".join(m.groups()[:6]) + f", {1 if m.group(7) else 0}" + f', "{plan}"}},')
    return "\n".join(rows) + "\n"


def _write_if_changed(path: str, text: str) -> None:
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)


def build(verbose: bool = False, jobs: int | None = None) -> str:
    sys.path.insert(0, HERE)
    try:
        import fixed_plans
    finally:
        sys.path.pop(0)
    os.makedirs(OBJ, exist_ok=True)
    _write_if_changed(os.path.join(OBJ, "shipped_plans.inc"), shipped_plans_inc())
    shapes = fixed_plans.shipped_shapes(extra=os.environ.get("WSIM_FIXED_EXTRA", ""))
    units = [(os.path.join(CSRC, "mbs.hip"), os.path.join(OBJ, "mbs.o"), ["-I", OBJ]),
             (os.path.join(CSRC, "problem.hip"), os.path.join(OBJ, "problem.o"), [])]
    for ct, m, p, e in VARIANTS:
        for c in COMPONENTS:
            for k in (0, 1) if e in PSTD_EPILOGUES else (0, 1, FAMILY_F64):
                name = f"v_{ct}_{m}_{p}_{e}_{c}_{k}"
                units.append((os.path.join(CSRC, "variant.hip"), os.path.join(OBJ, name + ".o"),
                              [f"-DWSIM_CONTIG={ct}", f"-DWSIM_MODE={m}", f"-DWSIM_PRO={p}",
                               f"-DWSIM_EPI={e}", f"-DWSIM_C={c}", f"-DWSIM_COMPACT={k}"]))
    for i, shape in enumerate(shapes):
        src = os.path.join(OBJ, f"fixed_{i}.hip")
        _write_if_changed(src, fixed_plans.kernel_unit(i, shape))
        units.append((src, os.path.join(OBJ, f"fixed_{i}.o"), []))
    for f in os.listdir(OBJ):  # units of shapes no longer generated
        m = re.fullmatch(r"fixed_(\d+)\.(hip|o)", f)
        if m and int(m.group(1)) >= len(shapes):
            os.remove(os.path.join(OBJ, f))
    table = os.path.join(OBJ, "fixed_table.hip")
    _write_if_changed(table, fixed_plans.table_unit(shapes))
    units.append((table, os.path.join(OBJ, "fixed_table.o"), []))
    fno_obj = os.path.join(OBJ, "fno.o")
    jobs = jobs or min(16, os.cpu_count() or 4)
    cmds = [[hipcc()] + _flags() + defs + ["-c", src, "-o", obj] for src, obj, defs in units]
    cmds.append([hipcc()] + _flags() + ["-c", os.path.join(CSRC, "fno.hip"), "-o", fno_obj])
    with cf.ThreadPoolExecutor(jobs) as ex:
        for cmd, msg in zip(cmds, ex.map(_run, cmds)):
            if verbose:
                print(cmd[-1], flush=True)
                if msg.strip():
                    print(msg)
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [u[1] for u in units])
    os.replace(tmp, LIB)
    tmp = FNO_LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, fno_obj])
    os.replace(tmp, FNO_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
    print(LIB)
