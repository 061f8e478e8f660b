"""bench.py's headline roofline reads the committed counter summary of the
kernel it names, and its flop counts are the stated per-unit figures."""
import json
import os

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_counter_summary_names_the_headline_kernel():
    name = bench.site_kernel_name({"lambda3": 0.0}, 4, 1, 512)
    c = bench.read_counters(name)
    assert c is not None, name
    assert abs(c["bytes_per_launch"] - c["fetch_bytes"] - c["write_bytes"]) <= 1
    assert c["waves"] == 32768 * 4 // 64
    # the traffic is within a small factor of the algorithmic bytes
    alg = c["algorithmic_bytes"]["read"] + c["algorithmic_bytes"]["write"]
    assert 1.0 <= c["bytes_per_launch"] / alg < 2.0
    assert bench.read_counters("k_tersoff_site<4, true>") is None


def test_flop_counts_per_unit():
    assert bench.tersoff_flops(1, 0, 0.0) == 45 + 1 + 18 + 48 + 8 * 20
    assert bench.tersoff_flops(0, 1, 0.0) == 160
    assert bench.tersoff_flops(0, 1, 1.7) == 160 + 9 + 7 * 20
    # diamond Si, 32768 atoms: 2 pairs and 12 ordered triplets per atom
    assert bench.tersoff_flops(65536, 393216, 0.0) / 32768 == 2464
    assert bench.sw_flops(1, 0) == 81 + 60 and bench.sw_flops(0, 1) == 100 + 40
    assert bench.eam_flops(1, 0) == 46 + 2 * (40 + 30) + 80 + 80 + 20 + 33 and bench.eam_flops(0, 1) == 60


def test_host_description():
    h = bench.host_cpu()
    assert h["threads_all_cores"] >= 1 and h["physical_cores_allowed"] >= 1
    assert isinstance(h["cpu_model"], str)
