"""Host-side checks of the profiling plumbing (no GPU)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_pass_names_from_template_arguments():
    from wavesim_amd.rocprof_summary import short_name
    assert short_name("void wsim::axis_pass_kernel<true, 1, 0, 1, 1, 256, 32>(wsim::PassArgs const*)") \
        == "pass_inv_mix[C=1]"
    assert short_name("void wsim::axis_pass_kernel<false, 2, 0, 0, 3, 256, 24>(x)") == "pass_mid_green[C=3]"
    assert short_name("void wsim::axis_pass_kernel<true, 1, 0, 3, 1, 512, 16>(wsim::PassArgs const*)") \
        == "pass_inv_mix_fwd[C=1]"
    assert short_name("void wsim::axis_pass_kernel<false, 2, 0, 0, 1, 320, 16, float, "
                      "wsim::FixedPlan<320, 16, 320, 3, false, 20, 16> >(wsim::PassArgs const*)") \
        == "pass_mid_green[C=1,fixed]"
    assert short_name("wsim::control_kernel(wsim::DevState*, double const*, int)") == "control"


def test_measured_traffic_matches_algorithmic_bytes():
    """The committed counter measurement of the 320^3 workload: each pass moves
    its algorithmic bytes (within 1 %), i.e. no pass re-reads HBM."""
    b = _bench()
    plan = None
    for line in open(os.path.join(ROOT, "profiles", "hbm_traffic_disordered_3d_256.txt")):
        if line.startswith("plan "):
            plan = line.split()[1]
    t = b.load_traffic("disordered_3d_256", plan)
    assert set(t) == {"pass_fwd_pro", "pass_fwd", "pass_mid_green", "pass_inv", "pass_inv_mix_fwd"}
    nvox = 320 ** 3
    for name in ("pass_fwd", "pass_mid_green", "pass_inv"):
        assert abs(sum(t[name]) / (16 * nvox) - 1) < 0.01
    # T, dE', gamma in; dE', next T out (40 B/voxel), E' in and out on the 12 even of the 24 traced
    # iterations (+8 B/voxel on average), the source in iterations 1-7 (Nw = 8; +8 B x 7/24)
    assert abs(sum(t["pass_inv_mix_fwd"]) / ((48 + 8 * 7 / 24) * nvox) - 1) < 0.01
    assert b.load_traffic("disordered_3d_256", "another-plan") == {}


def test_profiles_measured_on_the_shipped_plan():
    """bench.py reports the counter traffic only for the plan it was measured
    with, so the committed measurement must use the plan that ships for the
    320^3 workload."""
    import json
    shipped = json.load(open(os.path.join(ROOT, "wavesim_amd", "tuned_plans.json")))
    plan = shipped["320x320x320x1|rhs1|nw8"]["plan"]
    assert _bench().load_traffic("disordered_3d_256", plan)
    head = open(os.path.join(ROOT, "profiles", "kernel_stats_disordered_3d_256.txt")).read(2000)
    assert plan in head
