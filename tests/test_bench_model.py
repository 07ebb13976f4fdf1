"""CPU checks of the benchmark's measurement model: the algorithmic byte
counts behind roofline.achieved and the committed per-variant HBM traffic
that roofline.traffic is read from."""
import json
import os

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("N", [256, 1024, 4096])
@pytest.mark.parametrize("s", [4, 8])
def test_algo_bytes_closed_forms(N, s):
    """Row pass 6 s N^2 (2 planes of M x N/2 complex in and out, M = 3N/2);
    column pass of the RK3 stages of one step 6 s N^2 for the product and
    velocity planes plus 7 omega / fnm1-sized planes (s N^2 each) over the
    three stages: 25/3 s N^2 on average; CNAB2 reads and writes omega and
    fnm1: 10 s N^2."""
    row, col = bench.algo_bytes(N, s)
    assert row == 6 * s * N * N
    # the stage average is held in whole bytes: 7 s N^2 / 3 rounded down
    assert col == 6 * s * N * N + (7 * s * N * N) // 3
    assert -3 < 3 * col - 25 * s * N * N <= 0
    row_c, col_c = bench.algo_bytes(N, s, method="CNAB2")
    assert (row_c, col_c) == (6 * s * N * N, 10 * s * N * N)


def test_traffic_file_covers_every_variant_with_sane_ratios():
    """Every compiled step-pass variant of the N=4096 plans has a measured
    traffic entry (both passes, both precisions), each from three launches,
    and measured bytes are at least the compulsory bytes (less would mean a
    mis-scaled counter)."""
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        t = json.load(f)
    k = t["kernels"]
    assert abs(t["fetch_scale"] - 2.0) < 0.01 and abs(t["write_scale"] - 1.0) < 0.01
    for real, s in (("float", 4), ("double", 8)):
        row_b, col_b = bench.algo_bytes(4096, s)
        for v in range(8):
            for name, algo in ((f"k_row<{real},6144,STEP,V{v}>", row_b), (f"k_col<{real},6144,RK3,V{v}>", col_b)):
                e = k[name]
                assert e["algo_bytes"] == algo
                assert len(e["fetch_kib"]) == 3 and len(e["write_kib"]) == 3
                assert e["bytes"] >= 0.99 * algo
                assert abs(e["ratio"] - e["bytes"] / algo) < 1e-3


def test_traffic_lookup_by_kernel_name():
    t = bench.load_traffic()
    row_b, col_b = bench.algo_bytes(4096, 4)
    b, src = bench._traffic_bytes(t, "k_col<float,6144,RK3,V3>", 4096, col_b)
    assert b == t["kernels"]["k_col<float,6144,RK3,V3>"]["bytes"] and "traffic-sweep" in src
    assert bench._traffic_bytes(t, "k_col<float,6144,RK3,V3>", 2048, col_b) == (None, None)
    assert bench._traffic_bytes(t, "k_col<float,6144,RK3,V9>", 4096, col_b) == (None, None)
