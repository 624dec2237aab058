"""The rocprofv3 summaries of tlaloque_amd/profiling.py on small synthetic
CSV files: the traffic calibration on the plain copy kernel (not its
variants), the per-unit scaling of the counters report, and the kernel
statistics.  Host code only."""
import csv
import json

from tlaloque_amd import profiling

PMC_FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _pmc(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=PMC_FIELDS)
        w.writeheader()
        for i, (k, c, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v})


def test_traffic_calibrates_on_the_plain_copy_kernel(tmp_path):
    copy = 1 << 20
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    # the plain copy counts half the bytes it reads, a nontemporal variant all of them
    _pmc(f, [("k_stream_copy(float4 const*, float4*, long long)", "FETCH_SIZE", copy / 2048),
             ("void k_stream_copy_x<2, true>(float const*, float*, long long)", "FETCH_SIZE",
              copy / 1024),
             ("void tlq::k_stage2<float, 64, 8, 64, 2, 0, 1>(tlq::Problem const*)", "FETCH_SIZE",
              3000.0),
             ("void tlq::k_stage2<float, 64, 8, 64, 2, 0, 1>(tlq::Problem const*)", "FETCH_SIZE",
              1000.0)])
    _pmc(w, [("k_stream_copy(float4 const*, float4*, long long)", "WRITE_SIZE", copy / 1024),
             ("void k_stream_copy_x<2, true>(float const*, float*, long long)", "WRITE_SIZE",
              copy / 1024),
             ("void tlq::k_stage2<float, 64, 8, 64, 2, 0, 1>(tlq::Problem const*)", "WRITE_SIZE",
              500.0)])
    rep = profiling.traffic_report(str(f), str(w), copy, 1000, "f", "w", "t")
    js = json.loads([ln for ln in rep.splitlines() if ln.startswith("traffic-json: ")][0][14:])
    assert js["calibration"]["fetch_factor"] == 2.0
    assert js["calibration"]["write_factor"] == 1.0
    k = js["kernels"]["k_stage2"]
    # mean of the two fetch launches (2000 KiB) x 2, plus 500 KiB written
    assert k["read_bytes"] == 2000 * 1024 * 2.0 and k["write_bytes"] == 500 * 1024
    assert k["bytes_per_zone"] == (4000 + 500) * 1024 / 1000


def test_counters_scale_wave_counts_to_lanes(tmp_path):
    p = tmp_path / "p.csv"
    kern = "void tlq::k_cool_dense<float, 2>(tlq::Problem const*)"
    _pmc(p, [(kern, "SQ_INSTS_VALU", 1000.0), (kern, "SQ_INSTS_VALU", 3000.0),
             (kern, "SQ_ACTIVE_INST_VALU", 2500.0), (kern, "SQ_BUSY_CYCLES", 800.0),
             ("void tlq::k_stage1<float, 32, 8, 32, 0, 3>(tlq::Problem const*)", "SQ_INSTS_VALU",
              9.0)])
    rep = profiling.counters_report([str(p)], "k_cool_dense", 128.0, "cells", "cmd", "t")
    rows = {ln.split("|")[2].strip(): ln.split("|")[3:6] for ln in rep.splitlines()
            if ln.startswith("| `tlq::k_cool_dense")}
    assert "k_stage1" not in rep
    mean, per_unit, n = (x.strip() for x in rows["SQ_INSTS_VALU"])
    assert float(mean) == 2000.0 and float(per_unit) == 2000.0 * 64 / 128 and n == "2"
    assert float(rows["SQ_ACTIVE_INST_VALU"][1]) == 2500.0 * 64 / 128
    assert float(rows["SQ_BUSY_CYCLES"][1]) == 800.0 / 128  # not a per-wave count


def test_kernel_stats_per_kernel(tmp_path):
    t = tmp_path / "t.csv"
    fields = ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "VGPR_Count", "Accum_VGPR_Count",
              "SGPR_Count", "LDS_Block_Size", "Scratch_Size", "Grid_Size", "Workgroup_Size"]
    with open(t, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=fields)
        w.writeheader()
        for s, e in ((0, 2000), (5000, 9000)):
            w.writerow({"Kernel_Name": "void tlq::k_stage1<float, 32, 8, 32, 0, 3>(tlq::Problem const*)",
                        "Start_Timestamp": s, "End_Timestamp": e, "VGPR_Count": 84,
                        "Accum_VGPR_Count": 0, "SGPR_Count": 112, "LDS_Block_Size": 26112,
                        "Scratch_Size": 80, "Grid_Size": 4194304, "Workgroup_Size": 256})
    st = {k["name"]: k for k in profiling.kernel_stats(str(t))}
    k = st["tlq::k_stage1<float, 32, 8, 32, 0, 3>"]
    assert k["calls"] == 2 and k["total"] == 6.0 and k["min"] == 2.0 and k["max"] == 4.0
