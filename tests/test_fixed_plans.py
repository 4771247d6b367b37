"""Launch shapes of the shipped plans (wavesim_amd/fixed_plans.py), which the
build compiles as plan-specialised kernels: the restatement of the host's tile
choice, checked on shapes worked out by hand from the LDS sizes."""
from __future__ import annotations

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "wavesim_amd"))
import fixed_plans as fp  # noqa: E402


def test_320_cubed_shipped_plan_shapes():
    """320^3 plan y:4,16,5/32/512;x:5,16,4/16;z:16,5,4/0.
    Fused y pass: 23 lines (23 x 361 x 8 B tile + 1288 B columns + 2 ramps of
    2560 B + 1280 B coordinates + 638 twiddles x 8 B = 79 224 B <= 80 KB; 24
    lines exceed it).  Strided x and z: 16 columns (32 need 320 x 33 x 8 B =
    84 480 B of tile alone)."""
    shapes = fp.launch_shapes([320, 320, 320], 1, "y:4,16,5/32/512;x:5,16,4/16;z:16,5,4/0")
    got = {(s[0], s[1], s[3]): s for s in shapes}
    fused = got[(1, 1, 3)]
    assert fused[6:9] == (320, 23, 512) and fused[11] == (4, 16, 5)
    assert fused[9] == 4 and fused[10] == 1  # two workgroups per CU: 4 waves per SIMD, early loads
    for key, rad in (((0, 0, 0), (5, 16, 4)), ((0, 1, 0), (5, 16, 4)), ((0, 2, 0), (16, 5, 4))):
        s = got[key]
        assert s[6:9] == (320, 16, 512) and s[11] == rad
        assert s[10] == 0  # three workgroups per CU: loads after the tables
    b = 23 * fp.contig_pitch(320) * 8 + fp.lds_header(23, 320, 2 * fp.twiddles([4, 16, 5]), True)
    assert b == 79224
    assert 24 * fp.contig_pitch(320) * 8 + fp.lds_header(24, 320, 638, True) > 80 * 1024


def test_1152_squared_two_dimensional_shapes():
    """2-D: the first axis fused (and its once-per-solve first forward pass;
    the PSTD update fused alike, the same tile as the MBS one), the second the
    middle pass (no x passes)."""
    shapes = fp.launch_shapes([1152, 1152, 1], 1, "y:8,3,3,16/24/512/noxcd;x:16,3,3,8/16/0/xcd")
    assert [s[:4] for s in shapes] == [fp.FUSED_FIRST, fp.FIRST_FWD, fp.PSTD_FUSED, fp.MID_STRIDED]
    assert shapes[2][4:] == shapes[0][4:]
    # strided 1152-point tiles: 4 columns (1152 x 9 x 8 B = 82 944 B > 80 KB for 8)
    assert shapes[3][7] == 4


def test_single_stage_first_axis_uses_separate_mix_pass():
    shapes = fp.launch_shapes([16, 16, 1], 1, "y:16/0;x:16/0")
    assert shapes[0][:4] == fp.MIX_FIRST


def test_shipped_shapes_are_distinct_and_generate_units():
    shapes = fp.shipped_shapes()
    assert len(shapes) == len(set(shapes)) > 0
    src = fp.kernel_unit(0, shapes[0])
    assert "FixedPlan<" in src and "wsim_fixed_0" in src
    table = fp.table_unit(shapes)
    assert table.count("wsim_fixed_") == 2 * len(shapes)


def test_complex128_plans_of_their_own(tmp_path):
    """A "|f64" entry gives complex128 solves of that shape their own plan
    (and plan-specialised kernels of that plan); a shape without one runs its
    complex64 plan in complex128 too."""
    import json
    t = {"320x320x320x1|rhs1|nw8": {"plan": "y:4,16,5/32/512;x:16,20/16/512;z:20,16/16/320"},
         "320x320x320x1|rhs1|nw8|f64": {"plan": "y:8,8,5/0;x:16,20/16/512;z:4,16,5/8/256"},
         "192x192x192x1|rhs1|nw8": {"plan": "y:3,16,4/16/256;x:4,16,3/32/256;z:3,8,8/16/0"}}
    p = tmp_path / "plans.json"
    p.write_text(json.dumps(t))
    shapes = fp.shipped_shapes(str(p))
    mid = lambda f64: {(s[6], s[11]) for s in shapes if s[:2] == (0, 2) and (s[5] == 2) == f64}
    assert mid(False) == {(320, (20, 16)), (192, (3, 8, 8))}   # the z (middle) passes
    assert mid(True) == {(320, (4, 16, 5)), (192, (3, 8, 8))}


def test_plan_lookup_by_precision(tmp_path, monkeypatch):
    import json
    from wavesim_amd import plans
    cache = tmp_path / "cache.json"
    cache.write_text(json.dumps({"8x8x8x1|rhs1|nw8": {"plan": "a"}, "8x8x8x1|rhs1|nw8|f64": {"plan": "b"},
                                 "16x16x16x1|rhs1|nw8": {"plan": "c"}}))
    monkeypatch.setenv("WSIM_PLAN_CACHE", str(cache))
    assert plans.key([8, 8, 8, 1], 1, 8, f64=True) == "8x8x8x1|rhs1|nw8|f64"
    assert plans.lookup([8, 8, 8, 1], 1, 8) == "a"
    assert plans.lookup([8, 8, 8, 1], 1, 8, f64=True) == "b"
    assert plans.lookup([16, 16, 16, 1], 1, 8, f64=True) == "c"  # no complex128 plan: the complex64 one
