"""Launch shapes of the shipped FFT plans, for plan-specialised pass kernels.

The solver's pass kernel reads its FFT stages from the launch arguments at
run time (one kernel for every plan).  For the plans tuned on MI355X and
shipped in ``tuned_plans.json`` the build also compiles kernels with the
plan as compile-time constants (transform length, radices, lines per
workgroup, threads): divisions by the stage spans and row counts become
constant multiply-shifts and LDS offsets become instruction immediates.
The host launches such a kernel only when a launch has exactly its shape and
falls back to the run-time-plan kernel otherwise, so a shape computed here
that the host does not reproduce costs speed, never correctness.

``launch_shapes`` restates the host's tile choice (``Solver::configure`` in
``csrc/mbs.hip``): the LDS bytes of a tile and its tables against the 80 KB
per-workgroup target, the per-stage register budget of the kernel family,
and the plan's caps on lines and threads.
"""
from __future__ import annotations

import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LDS_TARGET = 80 * 1024
LDS_DEVICE = 160 * 1024
PAD_SHIFT = 3
MAX_ITS = 6
# pass variants as (contig, mode, prologue, epilogue): mode 0 fwd, 1 inv, 2 mid
FWD_STRIDED, INV_STRIDED, MID_STRIDED, FUSED_FIRST, MIX_FIRST = (0, 0, 0, 0), (0, 1, 0, 0), (0, 2, 0, 0), (1, 1, 0, 3), (1, 1, 0, 1)
FIRST_FWD = (1, 0, 1, 0)  # the first forward pass (source, ramps): once per solve when the last pass is fused
PSTD_FUSED = (1, 1, 0, 5)  # the PSTD update fused with the next step's first forward pass (complex64 scalar)


def ept_for(C: int, family: int) -> int:
    return (16 if family else 32) if C == 1 else (12 if family else 24)


def its_per_stage(ept: int, C: int, R: int) -> int:
    q = ept // (C * R)
    return 1 if q < 1 else min(q, MAX_ITS)


def twiddles(rad) -> int:
    ns, t = 1, 0
    for R in rad:
        t += (R - 1) * ns
        ns *= R
    return max(t, 1)


def tile_fits(rad, n, ncol, C, family, nthr) -> bool:
    for R in rad:
        need = -(-ncol * (n // R) // nthr)
        if need > its_per_stage(ept_for(C, family), C, R):
            return False
    return True


def r16(x: int) -> int:
    return (x + 15) & ~15


def lds_header(ncol, n, ntw, fuse, f64=False) -> int:
    """Column table, ramps, coordinates and twiddles in front of the tile (complex64:
    56-byte column entries, 8-byte complex, 4-byte real; complex128: 96, 16, 8)."""
    ci, cz, rz = (96, 16, 8) if f64 else (56, 8, 4)
    return r16(ncol * ci) + (2 if fuse else 1) * r16(n * cz) + r16(n * rz) + ntw * cz


def contig_pitch(n: int) -> int:
    return n + (n >> PAD_SHIFT) + 1


def parse_plan(text: str):
    """{axis: (radices, ncol cap, threads)} of a plan string."""
    out = {}
    for item in text.split(";"):
        a = "yxz".index(item[0])
        parts = item[2:].split("/")
        rad = [int(x) for x in parts[0].split(",")]
        ncol = int(parts[1]) if len(parts) > 1 else 0
        thr = int(parts[2]) if len(parts) > 2 else 0
        out[a] = (rad, ncol, thr)
    return out


def configure(N, C, rad, ncap, thr, axis, contig, mode, fuse, f64=False):
    """(ncol, threads, family, waves, early) of one pass launch, as the host
    chooses them (family 0/1: complex64 kernels, 2: complex128)."""
    cz = 16 if f64 else 8
    base = 2 if f64 else 0
    n = N[axis]
    nthr = thr if thr > 0 else (256 if contig else 512)
    ntwf, ntwi = twiddles(rad), twiddles(rad[::-1])
    ntw = (0 if (mode == 1 and not fuse) else ntwf) + (0 if mode == 0 else ntwi)
    inner = 1
    for b in range(axis):
        inner *= N[b]
    nvox = N[0] * N[1] * N[2]

    def fits(c, cap):
        if not tile_fits(rad, n, c, C, base, nthr):
            return False
        b = C * c * contig_pitch(n) * cz if contig else n * (C * c + 1) * cz
        return b + lds_header(c, n, ntw, fuse, f64) <= cap

    cap_b = LDS_TARGET if fits(1, LDS_TARGET) else LDS_DEVICE
    cap = ncap if ncap > 0 else (16 if contig else 32)
    ncol = 0
    if contig:
        for c in range(cap, 0, -1):
            if fits(c, cap_b):
                ncol = c
                break
        ncol = min(ncol, nvox // n)
    else:
        c0 = 1
        while c0 * 2 <= cap:
            c0 *= 2
        c = c0
        while c >= 1:
            if fits(c, cap_b):
                ncol = c
                break
            c >>= 1
        while ncol > 1 and ncol // 2 >= inner:
            ncol >>= 1
    if ncol == 0:
        return None
    family = 2 if f64 else (1 if tile_fits(rad, n, ncol, C, 1, nthr) else 0)
    b = C * ncol * contig_pitch(n) * cz if contig else n * (C * ncol + 1) * cz
    lds = b + lds_header(ncol, n, ntw, fuse, f64)
    # waves per SIMD that the LDS allows: the kernel's register count must
    # not be what holds occupancy below it (4 SIMDs; at most 6 asked, so
    # that no kernel is squeezed below 80 registers and spills; complex128
    # kernels, whose values take twice the registers, at most 2)
    waves = min(2 if f64 else 6, max(1, (LDS_DEVICE // lds) * (nthr // 64) // 4))
    # first-stage loads before the table set-up where at most two workgroups
    # share a CU (measured on the 320^3 passes: the fused pass with two
    # workgroups per CU 326 -> 310 us, the strided passes with three slower)
    early = LDS_DEVICE // lds <= 2
    return ncol, nthr, family, waves, early


def launch_shapes(N, C, plan_text, f64=False):
    """The per-iteration pass launches of a complex64 solve on grid N with the
    given plan: [(contig, mode, pro, epi, C, family, n, ncol, threads, waves per SIMD, early loads,
    radices)]."""
    plan = parse_plan(plan_text)
    axes = [a for a in range(3) if N[a] > 1]
    if len(axes) < 2:
        return []
    out = []

    def add(variant, axis):
        rad, ncap, thr = plan[axis]
        contig, mode, pro, epi = variant
        cfg = configure(N, C, rad, ncap, thr, axis, bool(contig), mode, epi in (3, 5), f64)
        if cfg is None:
            return
        ncol, nthr, family, waves, early = cfg
        # a fixed plan's butterflies per thread stay within the family's budget
        if not all(-(-ncol * (N[axis] // R) // nthr) <= its_per_stage(ept_for(C, family), C, R) for R in rad):
            return
        out.append((contig, mode, pro, epi, C, family, N[axis], ncol, nthr, waves, int(early), tuple(rad)))

    first = axes[0]
    add(FUSED_FIRST if len(plan[first][0]) >= 2 else MIX_FIRST, first)
    add(FIRST_FWD, first)
    # a PSTD solver of the grid runs the same plan: its fused update (the other
    # PSTD passes are the MBS shapes)
    if C == 1 and not f64 and len(plan[first][0]) >= 2:
        add(PSTD_FUSED, first)
    for a in axes[1:-1]:
        add(FWD_STRIDED, a)
        add(INV_STRIDED, a)
    add(MID_STRIDED, axes[-1])
    return out


def shipped_shapes(path: str = os.path.join(HERE, "tuned_plans.json"), extra: str = ""):
    """Distinct launch shapes of every shipped plan (complex64), plus those of
    `extra` plans ("NyxNxxNzxC=plan", several joined by "|"; a build-time
    option for trying plans that need plan-specialised kernels, such as
    radix 20)."""
    with open(path) as f:
        table = json.load(f)
    items = {False: [], True: []}  # complex64 / complex128 plans
    for key in sorted(table):
        m = re.fullmatch(r"(\d+)x(\d+)x(\d+)x(\d+)\|rhs(\d+)\|nw(\d+)(\|f64)?", key)
        if m:
            items[m.group(7) is not None].append(([int(m.group(i)) for i in (1, 2, 3)], int(m.group(4)),
                                                  table[key]["plan"]))
        # a complex128 solve without a plan of its own runs the complex64 plan
        if m and m.group(7) is None and key + "|f64" not in table:
            items[True].append(items[False][-1])
    for item in filter(None, extra.split("|")):
        grid, plan = item.split("=", 1)
        n = [int(x) for x in grid.split("x")]
        items[False].append((n[:3], n[3], plan))
        items[True].append((n[:3], n[3], plan))
    shapes = []
    for f64 in (False, True):  # complex64, then the complex128 kernels
        for N, C, plan in items[f64]:
            for s in launch_shapes(N, C, plan, f64):
                if s not in shapes:
                    shapes.append(s)
    return shapes


def kernel_unit(i: int, shape) -> str:
    """C++ source of plan-specialised kernel number i (one compilation unit)."""
    ct, m, p, e, C, fam, n, ncol, nthr, waves, early, rad = shape
    rs = ", ".join(str(r) for r in rad)
    F = "double" if fam == 2 else "float"
    return ("// generated by wavesim_amd/build.py from tuned_plans.json (fixed_plans.py)\n"
            '#include "mbs_kernels.hpp"\n#include "mbs_config.hpp"\n\nnamespace wsim {\n'
            f"KernelFn wsim_fixed_{i}() {{ return axis_pass_kernel<{'true' if ct else 'false'}, {m}, {p}, {e}, {C}, "
            f"{nthr}, {ept_for(C, fam)}, {F}, FixedPlan<{n}, {ncol}, {nthr}, {waves}, {'true' if early else 'false'}, {rs}>>; }}\n"
            "}  // namespace wsim\n")


def table_unit(shapes) -> str:
    """C++ source of the lookup table of the plan-specialised kernels."""
    lines = ["// generated by wavesim_amd/build.py from tuned_plans.json (fixed_plans.py)",
             '#include "mbs_config.hpp"', "", "namespace wsim {"]
    lines += [f"KernelFn wsim_fixed_{i}();" for i in range(len(shapes))]
    # a host function's static table: the device pass of hipcc emits no
    # initialiser of host function addresses
    lines.append("const FixedKernel* fixed_kernel_table() {")
    lines.append("  static const FixedKernel t[] = {")
    for i, (ct, m, p, e, C, fam, n, ncol, nthr, _, _, rad) in enumerate(shapes):
        r = ", ".join(str(x) for x in rad)
        lines.append(f"      {{{ct}, {m}, {p}, {e}, {C}, {fam}, {n}, {ncol}, {nthr}, {len(rad)}, {{{r}}}, wsim_fixed_{i}}},")
    lines.append("      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, {0}, nullptr}};")
    lines.append("  return t;")
    lines.append("}")
    lines.append("}  // namespace wsim")
    return "\n".join(lines) + "\n"
