"""bench.py's command-line contract that holds without a GPU: more than one
GPU is refused with a message, and the copy calibration and CPU baseline
helpers report what the profiles and the JSON line use."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""


def test_cpu_baseline_reports_one_core_and_all_cores():
    import bench
    b = bench.cpu_baseline(0.2)
    assert b["cores"] == 1 and b["kind"] == "port" and b["unit"] == "updates/ns" and b["value"] > 0
    assert 1 <= b["all_cores"]["cores"] <= 16 and b["all_cores"]["value"] > 0
    assert "metropolis.rs:8-27" in b["sample"]


def test_traffic_file_matches_the_profile_summary():
    import json
    import bench
    t = bench.load_traffic()
    d = json.load(open(os.path.join(ROOT, "profiles", "ising_half_kernel_pt_L8-128_R32_traffic.json")))
    assert t == d["hbm_bytes_per_launch"]
    scaled = d["fetch_kib"] * 1024 * d["fetch_scale"] + d["write_kib"] * 1024 * d["write_scale"]
    assert abs(scaled - t) <= 1.0
    assert 0.99 < t / d["algorithmic_bytes_per_launch"] < 1.05


def test_a_step_is_one_exchange_interval_of_whole_graph_replays():
    import bench
    # every parallel-tempering workload exchanges once per step
    for name, (_, _, _, _, _, opts, _) in bench.WORKLOADS.items():
        assert opts["exchange_every"] == bench.SWEEPS_PER_STEP, name
        assert bench.SWEEPS_PER_STEP % opts["measure_every"] == 0, name
    for _, _, _, opts in bench.BATCH_MODELS:
        assert opts["exchange_every"] == bench.SWEEPS_PER_STEP
    # the timed sweeps are whole replays of one captured graph
    for steps in (1, 5, 20, 37, 300):
        sweeps = steps * bench.SWEEPS_PER_STEP
        g = bench.graph_length(sweeps)
        assert sweeps % g == 0 and 1 <= g <= 200
    assert bench.graph_length(200) == 200 and bench.graph_length(3000) == 200


def test_warmup_is_whole_replays_of_the_timed_graph():
    import bench
    for w, G, want in ((5, 200, 200), (0, 200, 200), (30, 200, 400), (20, 200, 200), (21, 200, 400), (3, 10, 30)):
        got = bench.warmup_sweeps(w, G)
        assert got == want and got % G == 0 and got >= w * bench.SWEEPS_PER_STEP


def test_learned_digest_keeps_the_ratios_of_each_model():
    import bench
    chain = {"equal_error_speedup_E": 1.2, "equal_error_speedup_U4": 0.9, "U4_err_mean": 0.004,
             "z_vs_plain_E": {"max_abs": 1.5, "frac_within_2": 1.0},
             "z_vs_plain_U4": {"max_abs": 2.5, "frac_within_2": 0.9}}
    res = {}
    for model in ("xy", "heisenberg"):
        r = {"sizes": [32], "temperatures": [1.0, 2.0], "plain": {"U4_err_mean": 0.005},
             "learned": dict(chain, learned_acceptance=0.99), "heat_bath": dict(chain)}
        for mode in ("learned", "heat_bath"):
            for q in ("E", "M2"):
                r[mode + "_over_plain_samples_" + q] = 1.25
        res[model] = r
    d = bench.learned_digest(res, 4000, 1000, 64, 400)
    assert d["sweeps"] == 4000 and d["warmup_sweeps"] == 1000 and d["hidden"] == 64
    for model in ("xy", "heisenberg"):
        m = d[model]
        assert m["learned_over_plain_samples_E"] == 1.25 and m["heat_bath_over_plain_samples_M2"] == 1.25
        assert m["learned_equal_error_speedup_U4"] == 0.9 and m["learned_z_vs_plain_max_U4"] == 2.5
        assert m["U4_err_mean"] == {"plain": 0.005, "learned": 0.004, "heat_bath": 0.004}
        assert m["learned_acceptance"] == 0.99
