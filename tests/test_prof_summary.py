"""The rocprofv3 CSV summaries (citcom_amd/prof_summary.py) on small
hand-made files with the tool's column names."""
import csv

from citcom_amd import prof_summary


def write(path, header, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def test_stats_trace_counters(tmp_path):
    s = tmp_path / "s.csv"
    write(s, ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"],
          [["void k_a<float>(Geo, float*)", 4, 4000, 1000, 80, 900, 1100, 10],
           ["k_b", 1, 1000, 1000, 20, 1000, 1000, 0]])
    out = prof_summary.stats(str(s), 10)
    lines = out.splitlines()
    assert lines[2].startswith("| `k_a<float>` | 4 | 0.00 | 1.00 | 80.0 % |")
    assert "5 launches" in out
    t = tmp_path / "t.csv"
    write(t, ["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z", "Start_Timestamp", "End_Timestamp"],
          [["k_a", 256, 1, 1, 0, 2000], ["k_a", 256, 1, 1, 10, 4010], ["k_a", 64, 1, 1, 0, 500],
           ["k_b", 64, 1, 1, 0, 100]])
    out = prof_summary.trace(str(t), "k_a")
    assert "| `k_a` | 256 | 2 | 3.00 | 0.01 |" in out and "| `k_a` | 64 | 1 | 0.50 | 0.00 |" in out
    assert "k_b" not in out
    c = tmp_path / "c.csv"
    write(c, ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"],
          [["k_a", 256, "FETCH_SIZE", 10], ["k_a", 256, "FETCH_SIZE", 30], ["k_a", 64, "FETCH_SIZE", 5]])
    out = prof_summary.counters(str(c), ".")
    assert "| `k_a` | 256 | FETCH_SIZE | 2 | 20.0 |" in out and "| `k_a` | 64 | FETCH_SIZE | 1 | 5.0 |" in out
