"""bench.py's contract pieces that need no GPU: the one-GPU guard and the
byte models its roofline divides by, against the figures the profile notes
quote, and the measured-traffic file's per-iteration / per-round split
against the two counter runs it was made from."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_algorithmic_bytes_of_the_headline_workload():
    b = _bench()
    tot, per_it, per_tr = b.algorithmic_bytes(256, 20000, 3, 5)
    assert round(per_it / 1e6, 2) == 10.07
    assert round(per_tr / 1e6, 2) == 6.07
    assert tot == 3 * per_it + 5 * per_tr


def test_toeplitz_byte_model():
    b = _bench()
    tot, per_it, per_rd = b.toeplitz_bytes(256, 3, 4)
    assert round(per_it / 1e6, 2) == 2.10
    assert round(per_rd / 1e6, 2) == 15.77
    assert tot == 3 * per_it + 4 * per_rd


def test_traffic_file_reproduces_its_counter_runs():
    with open(os.path.join(ROOT, "profiles", "hbm_traffic_toeplitz256_m20000.json")) as f:
        t = json.load(f)
    # the per-iteration / per-round figures reproduce each run's measured bytes
    # at that run's counts (read and write split by the model's ratios)
    b = _bench()
    P, T = 256 * 256, 256 * 257
    for kind, ratio_key in (("fetch", "fetch_ratio_to_model"), ("write", "write_ratio_to_model")):
        it, rd = t[kind + "_iterations"], t[kind + "_rounds"]
        model = ((64 * P + 80 * T) * rd + 16 * P * it) if kind == "fetch" else ((32 * P + 64 * T) * rd + 16 * P * it)
        assert abs(model * t[ratio_key] - t[kind + "_bytes"]) <= 1e-3 * t[kind + "_bytes"]
    both = t["bytes_per_iteration"] * t["fetch_iterations"] + t["bytes_per_round"] * t["fetch_rounds"]
    assert abs(both - (t["fetch_bytes"] + t["write_bytes"])) <= 0.01 * (t["fetch_bytes"] + t["write_bytes"])
    tot, _, _ = b.toeplitz_bytes(256, t["fetch_iterations"], t["fetch_rounds"])
    assert 0.8 < (t["fetch_bytes"] + t["write_bytes"]) / tot < 1.5
