"""CPU checks of the profile summaries behind the bench's traffic and kernel
figures (swd_amd/profile_fd.py), on hand-made rocprofv3 CSV files:
FETCH_SIZE / WRITE_SIZE scaled by the calibration copy's readings of a known
byte count, the counter rows of one dispatch summed over the dimensions they
are split into, and the trace summary keeping the last dispatches in start
order (the timed steps), not in file order."""
import csv
import json
import os

import pytest

from swd_amd import profile_fd as PF

ADJ = "void swd::fd_adj_tb<float, 5, false>(swd::FdJob<float> const*, swd::BlockMap const*, int, int)"
FWD = "void swd::fd_fwd_tb<float, 5, false>(swd::FdJob<float> const*, swd::BlockMap const*, int, int)"
COPY = "swd::calib_copy(HIP_vector_type<float, 4u> const*, HIP_vector_type<float, 4u>*, unsigned long)"


def write_counters(path, counter, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "1_counter_collection.csv"), "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for disp, kern, val in rows:
            w.writerow([disp, kern, counter, val])


def test_summarize_scales_by_the_copy_and_sums_dimensions(tmp_path, monkeypatch, capsys):
    monkeypatch.setattr(PF, "TRAFFIC_JSON", str(tmp_path / "traffic.json"))
    d = tmp_path / "prof"
    # the copy of 1 GiB reads back half its bytes on FETCH_SIZE (scale 2) and
    # all of them on WRITE_SIZE (scale 1); each adjoint dispatch's FETCH_SIZE
    # comes in two dimension rows (100 + 50 KiB), its WRITE_SIZE in one
    write_counters(str(d / "fetch"), "FETCH_SIZE",
                   [(1, COPY, 524288), (2, ADJ, 100), (2, ADJ, 50), (3, ADJ, 110), (3, ADJ, 40)])
    write_counters(str(d / "write"), "WRITE_SIZE", [(1, COPY, 1048576), (2, ADJ, 200), (3, ADJ, 200)])
    PF.main(["summarize", str(d), "--shapes", '{"fd_adj_tb": [5, 26]}'])
    capsys.readouterr()
    res = json.load(open(tmp_path / "traffic.json"))
    assert res["fetch_scale"] == pytest.approx(2.0) and res["write_scale"] == pytest.approx(1.0)
    k = res["kernels"]["fd_adj_tb"]
    assert k["dispatches"] == 2
    assert k["fetch_kib"] == pytest.approx(150.0) and k["write_kib"] == pytest.approx(200.0)
    assert k["bytes_per_launch"] == pytest.approx(1024 * (150 * 2 + 200))
    assert res["by_shape"]["fd_adj_tb:5x26"]["bytes_per_launch"] == pytest.approx(1024 * 500)


def test_summarize_needs_the_calibration_copy(tmp_path, monkeypatch):
    monkeypatch.setattr(PF, "TRAFFIC_JSON", str(tmp_path / "traffic.json"))
    d = tmp_path / "prof"
    write_counters(str(d / "fetch"), "FETCH_SIZE", [(2, ADJ, 100)])
    write_counters(str(d / "write"), "WRITE_SIZE", [(2, ADJ, 100)])
    with pytest.raises(RuntimeError):
        PF.main(["summarize", str(d)])
    assert not (tmp_path / "traffic.json").exists()


def test_trace_keeps_the_last_dispatches_in_start_order(tmp_path, capsys):
    d = tmp_path / "trace" / "host"
    os.makedirs(d)
    rows = []
    # 6 early adjoint dispatches of 200 us (autotuning), then 4 of 90 us;
    # forward: 3 of 70 us, then 2 of 60 us; written to the file in reverse
    t = 0
    for us in [200] * 6 + [90] * 4:
        rows.append((ADJ, t, t + us * 1000)); t += 1_000_000
    for us in [70] * 3 + [60] * 2:
        rows.append((FWD, t, t + us * 1000)); t += 1_000_000
    with open(d / "1_kernel_trace.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for r in reversed(rows):
            w.writerow(r)
    PF.main(["trace", str(tmp_path / "trace"), "--last-adj", "4", "--last-fwd", "2"])
    out = json.loads(capsys.readouterr().out)
    assert out["fd_adj_tb"] == {"dispatches": 4, "of": 10, "mean_us": 90.0, "min_us": 90.0, "max_us": 90.0}
    assert out["fd_fwd_tb"] == {"dispatches": 2, "of": 5, "mean_us": 60.0, "min_us": 60.0, "max_us": 60.0}
