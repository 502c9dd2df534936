"""sparc_flows_amd.rocprof_summary on small synthetic rocprofv3 CSV files."""
import io

from sparc_flows_amd import rocprof_summary as rs


def _write(path, header, rows):
    path.write_text(",".join(header) + "\n" + "\n".join(",".join(str(v) for v in r) for r in rows) + "\n")


def test_stats_by_grid_keeps_grid_sizes_apart(tmp_path):
    hdr = ["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z", "Start_Timestamp", "End_Timestamp"]
    rows = [("k_a", 256, 1, 1, 0, 1000), ("k_a", 256, 1, 1, 2000, 4000),
            ("k_a", 256, 16, 1, 5000, 13000), ("k_b", 64, 1, 1, 14000, 14500)]
    _write(tmp_path / "run_kernel_trace.csv", hdr, rows)
    out = io.StringIO()
    rs.stats_by_grid(str(tmp_path), out=out)
    lines = out.getvalue().splitlines()
    assert lines[0].startswith("k_a") and "grid 256x16x1 calls 1" in lines[0] and "avg_us 8.00" in lines[0]
    assert "grid 256x1x1 calls 2" in lines[1] and "avg_us 1.50" in lines[1] and "min_us 1.00" in lines[1]
    assert "k_b" in lines[2] and "avg_us 0.50" in lines[2]


def test_pmc_means_per_kernel_and_grid(tmp_path):
    hdr = ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
    rows = [("k_a", 1024, "FETCH_SIZE", 10.0), ("k_a", 1024, "FETCH_SIZE", 30.0),
            ("k_a", 16384, "FETCH_SIZE", 100.0)]
    _write(tmp_path / "run_counter_collection.csv", hdr, rows)
    out = io.StringIO()
    rs.pmc(str(tmp_path), out=out, by_grid=True)
    text = out.getvalue()
    assert "grid 16384" in text and "mean 100.0" in text
    assert "grid 1024" in text and "dispatches 2 mean 20.0 min 10.0 max 30.0" in text
    out = io.StringIO()
    rs.pmc(str(tmp_path), out=out)
    assert "dispatches 3 mean 46.7" in out.getvalue()
