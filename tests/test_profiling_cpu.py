"""The rocprofv3 summary helpers on synthetic counter files (no GPU)."""
import csv
import os

from wavefiber3d_amd import profiling


def _write(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                           "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], r)))


def test_traffic_per_unit_and_counter_table(tmp_path):
    gib = 1 << 30
    # copy: FETCH_SIZE counts half a KiB per KiB, WRITE_SIZE one
    _write(tmp_path / "fc", [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", gib / 2048)])
    _write(tmp_path / "wc", [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", gib / 1024)])
    # two transitions = phi + theta dispatches each; the counter is split
    # over two rows of one dispatch (summed)
    _write(tmp_path / "f", [(1, "phi_tile_kernel", "FETCH_SIZE", 50), (1, "phi_tile_kernel", "FETCH_SIZE", 50),
                            (2, "theta_mfma_kernel", "FETCH_SIZE", 300),
                            (3, "phi_tile_kernel", "FETCH_SIZE", 100),
                            (4, "theta_mfma_kernel", "FETCH_SIZE", 300),
                            (5, "leaf_f32_kernel", "FETCH_SIZE", 7000)])
    _write(tmp_path / "w", [(1, "phi_tile_kernel", "WRITE_SIZE", 10), (2, "theta_mfma_kernel", "WRITE_SIZE", 30),
                            (3, "phi_tile_kernel", "WRITE_SIZE", 10), (4, "theta_mfma_kernel", "WRITE_SIZE", 30),
                            (5, "leaf_f32_kernel", "WRITE_SIZE", 9000)])
    t = profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "fc"),
                          str(tmp_path / "wc"), float(gib), "phi_|theta_", units=2)
    assert t["dispatches"] == [4, 4]
    assert abs(t["fetch_bytes_per_kib"] - 2048.0) < 1e-9
    assert abs(t["write_bytes_per_kib"] - 1024.0) < 1e-9
    # (800 KiB x 2048 + 80 KiB x 1024) / 2 transitions
    assert abs(t["traffic_bytes_per_launch"] - (800 * 2048 + 80 * 1024) / 2) < 1e-6
    t1 = profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "fc"),
                           str(tmp_path / "wc"), float(gib), "theta_")
    assert abs(t1["traffic_bytes_per_launch"] - (300 * 2048 + 30 * 1024)) < 1e-6
    tab = profiling.counter_table([str(tmp_path / "f"), str(tmp_path / "w")], "phi_|leaf")
    lines = tab.splitlines()
    assert lines[0] == "| kernel | dispatches | FETCH_SIZE | WRITE_SIZE |"
    assert lines[2] == "| `leaf_f32_kernel` | 1 | 7000 | 9000 |"
    assert lines[3] == "| `phi_tile_kernel` | 2 | 100 | 10 |"
