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
    assert short_name("wsim::control_cycle_kernel(wsim::DevState*, double const*, int, int, long long)") \
        == "control_cycle"


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


def test_network_gemm_roles_follow_the_forward(monkeypatch):
    """The role list the trace summary labels a forward's GEMM launches with
    matches the launches FNO.forward makes (run on the CPU with the kernel
    library stubbed out, recording each GEMM's shape)."""
    import torch
    from wavesim_amd.learned import fno, kernels as K

    class Stub:
        def __getattr__(self, name):
            return lambda *a: 0

    calls = []
    monkeypatch.setattr(K, "lib", lambda: Stub())
    monkeypatch.setattr(K, "stream_ptr", lambda stream=None: None)
    monkeypatch.setattr(K, "gemm", lambda M, N, Kd, *a, **k: calls.append((M, N, Kd)))
    L, W, m = 8, 4, 2
    cfg = fno.Config(L=L, cin=6, width=W, modes=m, layers=2, proj=8)
    net = fno.FNO(cfg, device="cpu")
    net.forward(torch.zeros(1, 6, L ** 3), keep=False)
    roles = fno.forward_gemm_roles(cfg)
    assert len(calls) == len(roles) == 3 + 2 * 10
    shapes = {"lift": (W, L ** 3, 6), "pointwise Wp": (W, L ** 3, W), "complex mixing": (2 * W, L * L, 2 * W),
              "truncated DFT x": (2 * m, W * L * L, L), "inverse truncated DFT x": (W * L * L, L, 2 * m),
              "truncated DFT y": (2 * m, L, L), "inverse truncated DFT y": (L, L, 2 * m),
              "truncated DFT z": (2 * m, L * L, L), "inverse truncated DFT z": (L, L * L, 2 * m),
              "projection Q1": (8, L ** 3, W), "projection Q2": (2, L ** 3, 8)}
    for role, shape in zip(roles, calls):
        assert shapes[role] == shape, role


def _write_csv(path, header, rows):
    import csv
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_summaries_leave_out_launches_queued_after_a_stop(tmp_path):
    """Pass launches that return at once (queued after the solve stopped) do
    not dilute the per-pass averages of the trace, counter and traffic
    summaries."""
    from wavesim_amd import rocprof_summary as rs
    k = "void wsim::axis_pass_kernel<true, 1, 0, 3, 1, 512, 16>(wsim::PassArgs const*)"
    t0 = 0
    rows = []
    for dur in [300_000] * 6 + [5_000] * 2:
        rows.append([k, t0, t0 + dur])
        t0 += dur + 1000
    _write_csv(str(tmp_path / "tr" / "x_kernel_trace.csv"), ["Kernel_Name", "Start_Timestamp", "End_Timestamp"], rows)
    line = [l for l in rs.stats(str(tmp_path / "tr")).splitlines() if l.startswith("pass_inv_mix_fwd")][0].split()
    assert line[1] == "6" and float(line[2]) == 300.0 and line[-1] == "2"
    crow = [[k, "SQ_INSTS_VALU", v] for v in [4e6] * 6 + [10.0] * 2]
    _write_csv(str(tmp_path / "sq" / "pmc_counter_collection.csv"), ["Kernel_Name", "Counter_Name", "Counter_Value"], crow)
    table = rs.counters([str(tmp_path / "sq")], str(tmp_path / "tr"))
    valu = [l for l in table.splitlines() if l.startswith("SQ_INSTS_VALU")][0].split()
    assert float(valu[4]) == 4.0  # the fused pass's column: 4 M per busy dispatch


def test_network_forward_summary_labels_the_last_forward(tmp_path):
    """The per-role table of a network forward reads the GEMM launches that
    follow the last forward's first launch (the mixing-weight assembly)."""
    from wavesim_amd import rocprof_summary as rs
    from wavesim_amd.learned.fno import Config, forward_gemm_roles
    roles = forward_gemm_roles(Config(layers=4))
    rows, t = [], 0
    for fwd in range(2):
        rows.append(["wfno::complex_weights_kernel(float const*)", t, t + 5000])
        t += 6000
        for k, role in enumerate(roles):
            dur = 10_000 if role == "complex mixing" else 20_000 + 1000 * fwd
            rows.append(["wfno::gemm_kernel(wfno_gemm_args)", t, t + dur])
            t += dur + 100
            if k % 10 == 0:
                rows.append(["wfno::bias_act_kernel(float*)", t, t + 3000])
                t += 3100
    _write_csv(str(tmp_path / "fno" / "x_kernel_trace.csv"), ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
               rows)
    out = rs.fno_forward(str(tmp_path / "fno"))
    table = {l.split()[0] + " " + l.split()[1]: l.split() for l in out.splitlines()[2:]}
    assert float(table["complex mixing"][3]) == 10.0 and table["complex mixing"][2] == "12"
    assert float(table["pointwise Wp"][3]) == 21.0  # the second (last) forward's durations
    assert f"{len(roles)} GEMM launches" in out.splitlines()[0]
