"""Trace decoder (fherma-bytecrt_amd/trace.py) on synthetic rocprofv3 CSV files."""
import csv
import importlib.util

import pytest

from conftest import PKG_DIR


@pytest.fixture(scope="module")
def tr():
    spec = importlib.util.spec_from_file_location("bcrt_trace", PKG_DIR / "trace.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_union_and_lane_stats(tr):
    assert tr.union([(5, 7), (0, 2), (1, 3), (7, 9), (10, 10)]) == [(0, 3), (5, 9)]
    s = tr.lane_stats([(0, 2), (1, 3), (5, 9)])
    assert s == {"count": 3, "busy_ns": 7, "coverage_ns": 9, "idle_ns": 2, "max_gap_ns": 2}
    c = tr.lane_stats([(0, 2), (1, 3), (5, 9)], 2, 6)
    assert c["busy_ns"] == 2 and c["head_idle_ns"] == 0 and c["tail_idle_ns"] == 0 and c["idle_ns"] == 2


def test_decode_synthetic(tr, tmp_path):
    d = tmp_path / "run" / "host" / "1"
    d.mkdir(parents=True)
    k = "bcrt::(anonymous namespace)::{}(bcrt::DevTables, unsigned int*)"
    write(d / "trace_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
          [[k.format("k_encode"), 100, 200], [k.format("k_middle"), 200, 260], [k.format("k_encode"), 400, 500],
           ["void at::native::foo(int)", 600, 610]])
    write(d / "trace_memory_copy_trace.csv", ["Direction", "Start_Timestamp", "End_Timestamp", "Bytes"],
          [["MEMORY_COPY_HOST_TO_DEVICE", 50, 150, 64], ["MEMORY_COPY_DEVICE_TO_HOST", 260, 300, 32],
           ["MEMORY_COPY_HOST_TO_DEVICE", 300, 420, 64]])
    write(d / "trace_marker_api_trace.csv", ["Function", "Thread_Id", "Start_Timestamp", "End_Timestamp"],
          [["bcrt:whole_call", 1, 0, 550], ["bcrt:pack", 1, 10, 50], ["bcrt:pack", 1, 250, 300],
           ["bcrt:unpack", 2, 300, 330], ["roctxOther", 1, 0, 1]])
    res = tr.decode(tmp_path / "run")
    assert res["kernels"]["k_encode"]["calls"] == 2 and res["kernels"]["k_middle"]["calls"] == 1
    assert abs(res["kernels"]["k_encode"]["share"] - 200 / 270) < 1e-12
    assert res["lanes"]["compute"]["busy_ns"] == 270
    assert res["lanes"]["h2d"]["busy_ns"] == 220 and res["copy_bytes"] == {"h2d": 128, "d2h": 32}
    # GPU union: [50, 300) + [300, 500) + [600, 610) -> 460
    assert res["lanes"]["gpu_union"]["busy_ns"] == 460
    assert set(res["host_ranges"]) == {"bcrt:whole_call", "bcrt:pack", "bcrt:unpack"}
    (call,) = res["calls"]
    assert call["wall_ns"] == 550 and call["compute"]["busy_ns"] == 260 and call["bcrt:pack"]["busy_ns"] == 90
    assert call["gpu_union"]["busy_ns"] == 450 and call["compute"]["tail_idle_ns"] == 50
    text = tr.markdown(res)
    assert "| k_encode | 2 |" in text and "Per whole call" in text


def test_decode_counters_scaled(tr, tmp_path):
    write(tmp_path / "x_counter_collection.csv", ["Kernel_Name", "Counter_Name", "Counter_Value"],
          [["k_crt(x)", "FETCH_SIZE", "10"], ["k_crt(x)", "FETCH_SIZE", "30"], ["k_crt(x)", "WRITE_SIZE", "4"]])
    res = tr.decode(tmp_path, {"FETCH_SIZE": 2.0})
    assert res["counters"]["k_crt"] == {"FETCH_SIZE": 40.0, "WRITE_SIZE": 4.0}


def test_short_kernel_names(tr):
    sk = tr.short_kernel
    assert sk("void bcrt::(anonymous namespace)::k_middle(bcrt::DevTables, unsigned int*)") == "k_middle"
    assert sk("void bcrt::(anonymous namespace)::k_crt<15, false>(bcrt::DevTables, unsigned int const*)") == \
        "k_crt<15,false>"
    assert sk("void (anonymous namespace)::k_overlap<3, 0>(unsigned int, unsigned int, int*)") == "k_overlap<3,0>"
    assert sk("void bcrt::(anonymous namespace)::k_inv_cols<8>(bcrt::DevTables, unsigned int*)") == "k_inv_cols<8>"
