"""rocprofv3 CSV summaries (cuda_hlt_amd.profiling) on small files written
in rocprofv3's column layout."""
import csv

import pytest

from cuda_hlt_amd import profiling as pf

COUNTER_COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
                "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
                "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp",
                "End_Timestamp"]

COPY = "void at::native::vectorized_elementwise_kernel<4, at::native::CUDAFunctorOnSelf_add<int>>(int)"


def write_counters(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, COUNTER_COLS)
        w.writeheader()
        for d, name, counter, value in rows:
            w.writerow({c: 0 for c in COUNTER_COLS} | {"Dispatch_Id": d, "Kernel_Name": name, "Counter_Name": counter,
                                                        "Counter_Value": value})


def test_short_names():
    assert pf.short_name("velo::search_by_triplet_kernel(velo_search_by_triplet_args)") == "search_by_triplet_kernel"
    assert pf.short_name("void velo::consolidate_kernel<false>(velo_consolidate_velo_tracks_args)") == \
        "consolidate_kernel<false>"


def test_traffic_is_calibrated_on_the_copy_kernel(tmp_path):
    fetch, write = tmp_path / "f", tmp_path / "w"
    fetch.mkdir()
    write.mkdir()
    # the copy moves 1 MiB each way but FETCH_SIZE reports half of it
    write_counters(fetch / "run_counter_collection.csv", [
        (1, COPY, "FETCH_SIZE", 512.0),
        (2, "velo::search_by_triplet_kernel(a)", "FETCH_SIZE", 100.0),
        (3, "velo::search_by_triplet_kernel(a)", "FETCH_SIZE", 300.0),
        (4, "void velo::consolidate_kernel<false>(a)", "FETCH_SIZE", 10.0)])
    write_counters(write / "run_counter_collection.csv", [
        (1, COPY, "WRITE_SIZE", 1024.0),
        (2, "velo::search_by_triplet_kernel(a)", "WRITE_SIZE", 50.0),
        (4, "void velo::consolidate_kernel<false>(a)", "WRITE_SIZE", 20.0)])
    t = pf.traffic(fetch, write, copy_bytes=1 << 20)
    assert t["copy_kernel"]["fetch_factor"] == pytest.approx(2.0)
    assert t["copy_kernel"]["write_factor"] == pytest.approx(1.0)
    s = t["kernels"]["search_by_triplet_kernel"]
    assert s["fetch_bytes"] == 2 * 200 * 1024 and s["write_bytes"] == 50 * 1024   # mean of two dispatches
    assert t["kernels"]["consolidate_kernel<false>"]["fetch_bytes"] == 2 * 10 * 1024
    assert "at::native" not in str(t["kernels"])


def test_counter_rows_of_one_dispatch_are_summed(tmp_path):
    p = tmp_path / "run_counter_collection.csv"
    write_counters(p, [(7, "velo::k_kernel(a)", "SQ_WAVES", 3.0), (7, "velo::k_kernel(a)", "SQ_WAVES", 5.0),
                       (8, "velo::k_kernel(a)", "SQ_WAVES", 2.0)])
    c = pf.read_counters(p)
    assert c["k_kernel"]["SQ_WAVES"] == pytest.approx(5.0) and c["k_kernel"]["dispatches"] == 2


def test_trace_groups_by_launch_shape(tmp_path):
    p = tmp_path / "run_kernel_trace.csv"
    cols = ["Kernel_Name", "Grid_Size_X", "Workgroup_Size_X", "Start_Timestamp", "End_Timestamp"]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, cols)
        w.writeheader()
        w.writerow({"Kernel_Name": "velo::a_kernel(x)", "Grid_Size_X": 640, "Workgroup_Size_X": 64,
                    "Start_Timestamp": 0, "End_Timestamp": 2000})
        w.writerow({"Kernel_Name": "velo::a_kernel(x)", "Grid_Size_X": 640, "Workgroup_Size_X": 64,
                    "Start_Timestamp": 0, "End_Timestamp": 4000})
        w.writerow({"Kernel_Name": "velo::a_kernel(x)", "Grid_Size_X": 128, "Workgroup_Size_X": 64,
                    "Start_Timestamp": 0, "End_Timestamp": 1000})
    rows = pf.read_trace(p)
    assert rows[0] == {"kernel": "a_kernel", "workgroups": 10, "workgroup_size": 64, "dispatches": 2,
                       "average_us": 3.0, "total_us": 6.0}
    assert rows[1]["workgroups"] == 2 and rows[1]["average_us"] == 1.0


AGENT_COLS = ["Node_Id", "Agent_Type", "Cu_Count", "Simd_Count", "Num_Xcc", "Simd_Per_Cu", "Max_Waves_Per_Simd",
              "Lds_Size_In_Kb", "Max_Engine_Clk_Fcompute", "Name"]


def write_agents(path):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, AGENT_COLS)
        w.writeheader()
        w.writerow({"Node_Id": 0, "Agent_Type": "CPU", "Cu_Count": 16, "Simd_Count": 0, "Num_Xcc": 1,
                    "Simd_Per_Cu": 0, "Max_Waves_Per_Simd": 0, "Lds_Size_In_Kb": 0, "Max_Engine_Clk_Fcompute": 0,
                    "Name": "cpu"})
        w.writerow({"Node_Id": 1, "Agent_Type": "GPU", "Cu_Count": 256, "Simd_Count": 1024, "Num_Xcc": 8,
                    "Simd_Per_Cu": 4, "Max_Waves_Per_Simd": 8, "Lds_Size_In_Kb": 160,
                    "Max_Engine_Clk_Fcompute": 2400, "Name": "gfx950"})


def test_theoretical_occupancy_limits():
    agent = {"Simd_Per_Cu": 4, "Max_Waves_Per_Simd": 8, "Lds_Size_In_Kb": 160}
    # one-wave workgroups with 11 KiB of LDS: 14 fit in 160 KiB
    assert pf.theoretical_waves_per_cu(64, 64, 0, 11 * 1024, agent) == {"waves": 14, "max_waves": 32, "limit": "lds"}
    # 256 registers (VGPR + AGPR) leave two waves per SIMD
    assert pf.theoretical_waves_per_cu(256, 200, 56, 0, agent)["waves"] == 8
    assert pf.theoretical_waves_per_cu(256, 200, 56, 0, agent)["limit"] == "registers"
    # small kernels fill the wave slots
    assert pf.theoretical_waves_per_cu(1024, 32, 0, 0, agent) == {"waves": 32, "max_waves": 32,
                                                                 "limit": "wave slots"}


def test_lds_occupancy(tmp_path):
    write_agents(tmp_path / "run_agent_info.csv")
    rows = []
    k = "velo::search_by_triplet_kernel(a)"
    # 1 ms at 2.4 GHz = 2.4e6 cycles per XCD, summed over 8 XCDs in the row
    for d in (1, 2):
        for c, v in (("SQ_INSTS_LDS_LOAD_BANDWIDTH", 3e9 / 64), ("SQ_INSTS_LDS_STORE_BANDWIDTH", 1e9 / 64),
                     ("SQ_LDS_BANK_CONFLICT", 0.05 * 2.4e6 * 256), ("SQ_WAVE_CYCLES", 2.4e6 * 256 * 16 / 4),
                     ("GRBM_GUI_ACTIVE", 8 * 2.4e6)):
            rows.append({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v,
                         "Start_Timestamp": 0, "End_Timestamp": 1_000_000, "Workgroup_Size": 64,
                         "VGPR_Count": 64, "LDS_Block_Size": 11 * 1024})
    with open(tmp_path / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, COUNTER_COLS)
        w.writeheader()
        for r in rows:
            w.writerow({c: 0 for c in COUNTER_COLS} | r)
    t = pf.lds_occupancy(tmp_path)
    s = t["kernels"]["search_by_triplet_kernel"]
    assert s["dispatches"] == 2 and s["duration_us"] == pytest.approx(1000.0)
    assert s["lds_bytes"] == pytest.approx(4e9) and s["lds_TBps"] == pytest.approx(4.0)
    assert s["lds_bank_conflict_pct"] == pytest.approx(5.0)
    assert s["waves_per_cu"] == pytest.approx(16.0) and s["occupancy_pct"] == pytest.approx(50.0)
    assert "search_by_triplet_kernel" in pf.lds_markdown(t, "t", "c")
