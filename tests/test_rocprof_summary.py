"""The rocprofv3 summariser that writes the kernel tables and HBM traffic kept with the
measurements, on small hand-made traces: per-kernel aggregation and the blocked kernel's
per-launch-shape table of a kernel trace, and the copy-scaled FETCH_SIZE / WRITE_SIZE
bytes per k_pd_block launch."""

import csv

import rocprof_summary as R

FIELDS = ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X"]


def _trace(path, rows):
    path.mkdir()
    with open(path / "t_kernel_trace.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for name, t0, t1, grid in rows:
            w.writerow(dict(Kind="KERNEL_DISPATCH", Kernel_Name=name, Start_Timestamp=t0, End_Timestamp=t1,
                            Grid_Size_X=grid))


def test_short_name_strips_namespace_and_arguments():
    n = "void (anonymous namespace)::k_pd_block<22, 22, 5, 1024>(PDLevel, int, PDSched, PDConsts)"
    assert R.short_name(n) == "k_pd_block<22, 22, 5, 1024>"
    assert R.short_name("k_copy(float*, float const*, int)") == "k_copy"


def test_stats_aggregates_durations_and_launch_shapes(tmp_path):
    pd = "void (anonymous namespace)::k_pd_block<22, 22, 5, 1024>(PDLevel, int, PDSched, PDConsts)"
    rows = [(pd, 1000, 31000, 1401856), (pd, 40000, 74000, 1401856), (pd, 80000, 90000, 409600),
            ("k_derivatives(float2*, float const*)", 100000, 105000, 832)]
    d = tmp_path / "trace"
    _trace(d, rows)
    md = R.stats(str(d), "T", "CMD")
    lines = md.splitlines()
    assert lines[0] == "# T" and "CMD" in lines
    row = next(l for l in lines if l.startswith("| k_pd_block<22, 22, 5, 1024> | 3 |"))
    cells = [c.strip() for c in row.strip("|").split("|")]
    # 30 + 34 + 10 us over 3 calls; min 10, max 34; 74 of 79 us in total
    assert cells[1:] == ["3", "74", "24.667", "10.000", "34.000", "93.67"]
    assert "| k_pd_block<22, 22, 5, 1024> | 1401856 | 2 | 32.000 |" in lines
    assert "| k_pd_block<22, 22, 5, 1024> | 409600 | 1 | 10.000 |" in lines
    assert lines[-1] == "All k_pd_block launches: 3 calls, average 24.667 us."


def _counters(path, name, rows):
    path.mkdir()
    with open(path / "c_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for did, kname, grid, value in rows:
            w.writerow(dict(Dispatch_Id=did, Kernel_Name=kname, Grid_Size=grid, Counter_Name=name, Counter_Value=value))


def test_pmc_scales_by_the_copy_and_counts_each_level_second_warp(tmp_path):
    """Counters summed over a dispatch's rows, scaled so that the copy of 256 MiB reads
    256 MiB (FETCH_SIZE of the copy reads 128 MiB here: scale 2) and writes 256 MiB
    (scale 1); of the level's two warps (8 iterations, K = 4: two launches each) the
    second is the one counted, per launch."""
    copy, g = "k_copy(float*, float const*, long)", R.COPY_FLOATS // 4
    pd = "void (anonymous namespace)::k_pd_block<16, 16, 4, 1024>(PDLevel, int, PDSched, PDConsts)"
    _counters(tmp_path / "f", "FETCH_SIZE",
              [(1, copy, g, 65536.0), (1, copy, g, 65536.0), (2, pd, 1024, 99.0), (3, pd, 1024, 99.0),
               (4, pd, 1024, 10.0), (5, pd, 1024, 6.0), (5, pd, 1024, 6.0)])
    _counters(tmp_path / "w", "WRITE_SIZE",
              [(1, copy, g, 262144.0), (2, pd, 1024, 77.0), (3, pd, 1024, 77.0), (4, pd, 1024, 5.0),
               (5, pd, 1024, 7.0)])
    log = tmp_path / "probe.log"
    log.write_text("[probe] 16x16 tile (16, 16, 4, 1024)\n")
    r = R.pmc(str(tmp_path / "f"), str(tmp_path / "w"), str(log), 8)
    assert r["copy_kernel"]["fetch_scale"] == 2.0 and r["copy_kernel"]["write_scale"] == 1.0
    (lv,) = r["levels"]
    assert lv["size"] == [16, 16] and lv["launches_per_warp"] == 2
    assert lv["fetch_bytes"] == (10 + 12) * 1024 * 2 // 2
    assert lv["write_bytes"] == (5 + 7) * 1024 // 2
