"""bench.py's contract pieces that run without a GPU: the refusal of more than one
GPU, the per-launch byte formula, and the loaders of the committed profile figures."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and r.stdout == ""


def test_compulsory_bytes_by_hand():
    # 150+150 bases -> 5+5 words of 8 B; 44 B of metadata/results per pair; 4 B per run
    p = np.array([150, 33]); t = np.array([150, 1]); runs = np.array([7, 2])
    want = 8 * (5 + 5) + 8 * (2 + 1) + 2 * 44 + 4 * 9
    assert bench.compulsory_bytes(p, t, runs) == want


def test_profile_figures_load_only_for_their_workload():
    tr = bench.profile_json("traffic", "C4", 200000)
    assert tr is not None and tr["_path"].startswith("profiles/")
    # calibrated bytes = fetch KiB x factor + write KiB x factor
    want = 1024 * (tr["fetch_kib"] * tr["fetch_factor"] + tr["write_kib"] * tr["write_factor"])
    assert abs(tr["traffic_bytes"] - want) < 2
    assert tr["kernel"].startswith("bw::biwfa_persistent<")
    assert bench.profile_json("traffic", "C4", 20000) is None
    assert bench.profile_json("traffic", "C2", 1000000)["workload"] == "C2"
    assert bench.profile_json("work_counts", "C3", 10000)["workload"] == "C3"
    assert bench.profile_json("work_counts", "C3", 5000) is None
    # SQ counters: per-cell figures for the team kernel, per-pair ones for the LDS kernel
    sq4 = bench.profile_json("sq", "C4", 200000)
    assert abs(sq4["valu_per_cell"] - sq4["sq_insts_valu"] / sq4["cells_per_step"]) < 1e-3
    assert 0 < sq4["valu_per_cu_clock"] < 1 and sq4["kernel"] == tr["kernel"]
    assert "valu_per_pair" in bench.profile_json("sq", "C2", 1000000)
    assert bench.profile_json("sq", "C1", 10000) is None


def test_work_counts_are_consistent():
    for wl, n in (("C4", 200000), ("C2", 1000000)):
        c = bench.profile_json("work_counts", wl, n)
        assert c is not None
        assert abs(c["sum_ops"] - (11 * c["sum_cells"] + 4 * c["sum_words"])) < 1e-6 * c["sum_ops"]
        assert abs(c["ops_per_score_sq"] - c["sum_ops"] / c["sum_score_sq"]) < 1e-9 * c["ops_per_score_sq"]
        assert c["sample_pairs"] == len(range(0, n, c["every"])) and c["status_nonzero"] == 0
        # cells track score^2 closely: the ratio estimator's spread is tiny
        assert c["half_sample_ratio_spread"] < 0.01


def test_work_counts_match_the_restatement_on_a_subsample():
    from oracle.work_counts import count
    c = count("C2", every=1, threads=2, num_pairs=3000)
    assert c["sample_pairs"] == 3000
    full = bench.profile_json("work_counts", "C2", 1000000)
    # the first 3,000 pairs hold the same per-pair work as the whole set, within noise
    assert abs(c["ops_per_score_sq"] / full["ops_per_score_sq"] - 1) < 0.05


def test_host_cpu_reports_cores_and_model():
    h = bench.host_cpu()
    assert h["logical"] == os.cpu_count() and 1 <= h["usable"] <= h["logical"]
    assert h["physical"] is None or 1 <= h["physical"] <= h["logical"]


def test_gather_runs_picks_each_pairs_runs():
    class R:
        cigar_offset = np.array([0, 2, 3, 6], np.int64)
        cigar_runs = np.array([10, 11, 20, 30, 31, 32], np.uint32)
    off, runs = bench.gather_runs(R, np.array([0, 2], np.int64))
    assert off.tolist() == [0, 2, 5] and runs.tolist() == [10, 11, 30, 31, 32]
    assert bench._mismatched_pairs(off, runs, off, runs) == 0
    assert bench._mismatched_pairs(off, runs, off, np.array([10, 11, 30, 31, 33], np.uint32)) == 1
