"""CPU tests of bench.py's host logic: the one-GPU contract and the rocprof
summaries (kernel table split by stream, PMC traffic calibrated by the copy
kernel, the per-launch JSON the roofline reads).  No GPU is used."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, cwd=ROOT,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def write_csv(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_trace_summary_splits_streams(tmp_path):
    d = tmp_path / "c2"
    d.mkdir()
    write_csv(d / "c2_kernel_stats.csv",
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs"],
              [{"Name": "void cg::k_os_pass<8, 8, 16, false>(...)", "Calls": 3,
                "TotalDurationNs": 600000, "AverageNs": 200000, "Percentage": 100.0,
                "MinNs": 100000, "MaxNs": 300000}])
    trace = []
    for stream, name, dur in (("1", "cg::k_move_gen(...)", 50), ("1", "void cg::k_os_pass<8>(...)", 100),
                              ("1", "void cg::k_os_pass<8>(...)", 300), ("2", "void cg::k_os_pass<8>(...)", 200)):
        trace.append({"Kernel_Name": name, "Stream_Id": stream, "Start_Timestamp": 1000,
                      "End_Timestamp": 1000 + dur * 1000})
    write_csv(d / "c2_kernel_trace.csv", ["Kernel_Name", "Stream_Id", "Start_Timestamp", "End_Timestamp"],
              trace)
    r = run("--prof-summary", str(d))
    assert r.returncode == 0, r.stderr
    assert "| void cg::k_os_pass<8, 8, 16, false>(...) | 3 | 0.6 | 200.00 |" in r.stdout
    # main stream = the one running k_move: two passes averaging 200 us; the other one 200 us
    assert "| sort_pass | 2 | 200.00 | 1 | 200.00 |" in r.stdout


def test_pmc_summary_calibrates_and_splits_streams(tmp_path):
    fields = ["Dispatch_Id", "Queue_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    # a 1 GiB copy reads FETCH_SIZE 512 MiB and WRITE_SIZE 1 GiB (in KiB): factors 2 and 1
    rows_f = [{"Dispatch_Id": 1, "Queue_Id": 1, "Kernel_Name": "__amd_rocclr_copyBuffer",
               "Counter_Name": "FETCH_SIZE", "Counter_Value": 524288.0},
              {"Dispatch_Id": 2, "Queue_Id": 1, "Kernel_Name": "cg::k_move_gen(...)",
               "Counter_Name": "FETCH_SIZE", "Counter_Value": 1000.0},
              {"Dispatch_Id": 3, "Queue_Id": 1, "Kernel_Name": "void cg::k_os_pass<8>(...)",
               "Counter_Name": "FETCH_SIZE", "Counter_Value": 3000.0},
              {"Dispatch_Id": 4, "Queue_Id": 2, "Kernel_Name": "void cg::k_os_pass<8>(...)",
               "Counter_Name": "FETCH_SIZE", "Counter_Value": 1000.0}]
    rows_w = [{"Dispatch_Id": 1, "Queue_Id": 1, "Kernel_Name": "__amd_rocclr_copyBuffer",
               "Counter_Name": "WRITE_SIZE", "Counter_Value": 1048576.0},
              {"Dispatch_Id": 2, "Queue_Id": 1, "Kernel_Name": "cg::k_move_gen(...)",
               "Counter_Name": "WRITE_SIZE", "Counter_Value": 500.0},
              {"Dispatch_Id": 3, "Queue_Id": 1, "Kernel_Name": "void cg::k_os_pass<8>(...)",
               "Counter_Name": "WRITE_SIZE", "Counter_Value": 4000.0},
              {"Dispatch_Id": 4, "Queue_Id": 2, "Kernel_Name": "void cg::k_os_pass<8>(...)",
               "Counter_Name": "WRITE_SIZE", "Counter_Value": 2000.0}]
    for sub, rows in (("pmc_x_FETCH_SIZE", rows_f), ("pmc_x_WRITE_SIZE", rows_w)):
        (tmp_path / sub).mkdir()
        write_csv(tmp_path / sub / "c_counter_collection.csv", fields, rows)
    out = tmp_path / "traffic.json"
    r = run("--prof-summary", str(tmp_path), "--pmc", "--pmc-tag", "pmc_x", "--pmc-json", str(out),
            "--pmc-workload", "c2")
    assert r.returncode == 0, r.stderr
    js = json.loads(out.read_text())
    assert abs(js["calibration"]["fetch_factor"] - 2.0) < 1e-9
    assert abs(js["calibration"]["write_factor"] - 1.0) < 1e-9
    sp = js["kernels"]["sort_pass"]
    assert sp["launches"] == 2
    assert sp["read_bytes_per_launch"] == (3000 + 1000) * 1024 * 2 / 2
    assert sp["write_bytes_per_launch"] == (4000 + 2000) * 1024 / 2
    main = js["kernels_main_stream"]["sort_pass"]
    assert main["launches"] == 1
    assert main["read_bytes_per_launch"] == 3000 * 1024 * 2
    assert main["write_bytes_per_launch"] == 4000 * 1024
    assert js["workload"] == "c2"
