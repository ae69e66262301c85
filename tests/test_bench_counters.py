"""Host logic of the roofline's traffic figure: bench.counters_from turns two
rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE; csv) into bytes per launch,
scaled by the copy probe and averaged over the launches that ran, and
bench.dslash_traffic looks the autotuned launch up by kernel name and grid.
CPU only: the csv files are written here."""
import csv
import json
import os
import sys

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
          "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]
PROBE = "rhmc::k_copy_probe(HIP_vector_type<double, 4u> const*, HIP_vector_type<double, 4u>*, long)"
HOP = "void rhmc::k_dslash<float, 1, 4, 0, 2>(rhmc::DsArgs)"


def write_run(path, counter, rows):
    """rows: (kernel, grid, [per-dispatch values in KiB]); each value split
    over two csv lines of the same dispatch (the per-XCD instances)."""
    os.makedirs(path, exist_ok=True)
    did = 0
    with open(os.path.join(path, "x_counter_collection.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        for kernel, grid, values in rows:
            for v in values:
                did += 1
                for half in (0.25 * v, 0.75 * v):
                    w.writerow({"Correlation_Id": did, "Dispatch_Id": did, "Agent_Id": "Agent 1", "Queue_Id": 1,
                                "Process_Id": 1, "Thread_Id": 1, "Grid_Size": grid, "Kernel_Id": 1,
                                "Kernel_Name": kernel, "Workgroup_Size": 256, "LDS_Block_Size": 0,
                                "Scratch_Size": 0, "VGPR_Count": 8, "Accum_VGPR_Count": 0, "SGPR_Count": 16,
                                "Counter_Name": counter, "Counter_Value": half, "Start_Timestamp": 0,
                                "End_Timestamp": 1})


def test_counters_scaled_by_probe_and_noop_launches_left_out(tmp_path):
    gib_kib = float(1 << 20)
    # the probe reads as half its size on FETCH_SIZE, as its size on WRITE_SIZE
    fetch = [(PROBE, 2097152, [0.5 * gib_kib] * 3),
             (HOP, 262144, [60000.0, 59000.0, 61000.0, 50.0, 52.0])]
    write = [(PROBE, 2097152, [gib_kib] * 3),
             (HOP, 262144, [24000.0, 24000.0, 24000.0, 30.0])]
    write_run(str(tmp_path / "f"), "FETCH_SIZE", fetch)
    write_run(str(tmp_path / "w"), "WRITE_SIZE", write)
    out = bench.counters_from(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "c.json"))
    assert abs(out["factors"]["fetch"] - 2.0) < 1e-12 and abs(out["factors"]["write"] - 1.0) < 1e-12
    k = out["kernels"][f"{HOP}|262144"]
    # the two launches after the solve stopped (~50 KiB) are not averaged in
    assert k["dispatches"] == 3
    assert abs(k["read_bytes"] - 60000.0 * 1024 * 2.0) < 1e-6
    assert abs(k["write_bytes"] - 24000.0 * 1024) < 1e-6
    assert out["source_sha256"] == bench.kernel_source_hash()
    with open(tmp_path / "c.json") as fh:
        assert json.load(fh)["kernels"].keys() == out["kernels"].keys()


def test_traffic_lookup_of_the_autotuned_hop():
    """dslash_traffic finds the committed profile's row for the shape in a
    tuning table (one workgroup per tile for the hop) and refuses a shape
    that was not profiled."""
    table = [{"real_bytes": 4, "links": "fp32_12", "mode": 0, "reduction": 2, "nvec": 8, "rhs_per_thread": 4,
              "lanes_per_site": 1, "tile_order": "xcd", "tiles_per_workgroup": 1}]
    path = os.path.join(ROOT, "profiles", "dslash_32c8_counters.json")
    with open(path) as fh:
        prof = json.load(fh)
    t = bench.dslash_traffic(table, 32 * 32 * 32 * 8 // 2, 8, mode=0, red=2)
    if prof["source_sha256"] != bench.kernel_source_hash():
        assert t is None  # a profile of other kernel sources is not used
        return
    k = prof["kernels"][f"{HOP}|262144"]
    assert t == k["read_bytes"] + k["write_bytes"]
    table[0]["rhs_per_thread"] = 1  # a 1-per-lane hop was never profiled at 8 RHS
    assert bench.dslash_traffic(table, 32 * 32 * 32 * 8 // 2, 8, mode=0, red=2) is None


def test_bench_refuses_more_than_one_gpu():
    """One GPU only: --gpus above 1 exits with a message, before any GPU call."""
    import subprocess
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in json.loads(p.stdout.strip().splitlines()[-1])["error"]
