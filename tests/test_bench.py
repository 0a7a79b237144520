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


def test_algorithmic_bytes_by_hand():
    # 150+150 bases -> 5+5 words of 8 B; 44 B of metadata/results per pair; 4 B per run
    p = np.array([150, 33]); t = np.array([150, 1]); runs = np.array([7, 2])
    want = 8 * (5 + 5) + 8 * (2 + 1) + 2 * 44 + 4 * 9
    assert bench.algorithmic_bytes(p, t, runs) == want


def test_profile_figures_load_only_for_their_workload():
    with open(os.path.join(ROOT, "profiles", "c2_traffic.json")) as f:
        tr = json.load(f)
    got = bench.traffic_bytes("C2", 1000000)
    assert got == tr["traffic_bytes"]
    # calibrated bytes = fetch KiB x factor + write KiB x factor
    assert abs(got - 1024 * (tr["fetch_kib"] * tr["fetch_factor"] + tr["write_kib"] * tr["write_factor"])) < 2
    assert bench.traffic_bytes("C2", 1000) is None and bench.traffic_bytes("C4", 1000000) is None
    sq = bench.sq_counts("C2", 1000000)
    assert sq is not None
    assert abs(sq["salu_per_pair"] - sq["sq_insts_salu"] / sq["pairs"]) < 1e-6
    assert abs(sq["valu_per_pair"] - sq["sq_insts_valu"] / sq["pairs"]) < 1e-6
    assert bench.sq_counts("C4", 20000) is None
