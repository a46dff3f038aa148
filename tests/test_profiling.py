"""The rocprofv3 summary helpers on small hand-made CSVs: kernel names, the
largest-grid selection that separates batched launches from single-pair
ones, and the copy-kernel scaling of the traffic counters."""
import csv

from phaser_amd import profiling


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_ALL)
        w.writerow(header)
        w.writerows(rows)


def test_short_name_and_matches():
    n = profiling.short_name("void phz::(anonymous namespace)::k_fft_cols<-1, 16>(float2*, long)")
    assert n == "k_fft_cols<-1, 16>"
    assert profiling.short_name("phz::k_nn<4>(float4 const*, int)") == "k_nn<4>"
    assert profiling.matches("k_fft_cols<-1>", "k_fft_cols<-1, 16>")
    assert not profiling.matches("k_fft_cols<-1>", "k_fft_cols<1, 16>")
    assert profiling.matches("k_nn", "k_nn<4>")
    assert not profiling.matches("k_nn", "k_nnz")


def test_batched_stats_keep_largest_grid(tmp_path):
    head = ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X",
            "Grid_Size_Y", "Grid_Size_Z"]
    rows = [
        ["KERNEL_DISPATCH", "void phz::k_nn<4>(int)", 0, 1000, 256, 2, 1],   # one pair
        ["KERNEL_DISPATCH", "void phz::k_nn<4>(int)", 0, 9000, 256, 64, 1],  # batch
        ["KERNEL_DISPATCH", "void phz::k_nn<4>(int)", 0, 11000, 256, 64, 1],
        ["KERNEL_DISPATCH", "void phz::k_fuse(int)", 0, 500, 64, 1, 1],
    ]
    p = tmp_path / "trace.csv"
    _write(p, head, rows)
    st = profiling.batched_stats(str(p))
    g, d = st["k_nn<4>"]
    assert g == 256 * 64 and d == [9.0, 11.0]
    md = profiling.batched_stats_markdown(str(p))
    assert "| k_nn<4> | 16384 | 2 | 10.0 | 9.0 | 11.0 |" in md


def test_traffic_summary_scales_by_copy_kernel(tmp_path):
    head = ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
    copy = "void phz::(anonymous namespace)::k_copy_f4(float4*, float4 const*, unsigned long)"
    fetch = [[copy, 1000, "FETCH_SIZE", 512.0],
             ["void phz::k_nn<4>(int)", 64, "FETCH_SIZE", 100.0],
             ["void phz::k_nn<4>(int)", 4096, "FETCH_SIZE", 200.0],
             ["void phz::k_nn<4>(int)", 4096, "FETCH_SIZE", 300.0]]
    write = [[copy, 1000, "WRITE_SIZE", 1024.0],
             ["void phz::k_nn<4>(int)", 4096, "WRITE_SIZE", 10.0],
             ["void phz::k_nn<4>(int)", 4096, "WRITE_SIZE", 30.0]]
    pf, pw = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(pf, head, fetch)
    _write(pw, head, write)
    s = profiling.traffic_summary(str(pf), str(pw), 1 << 20)
    assert s["fetch_scale_bytes_per_unit"] == 2048.0
    assert s["write_scale_bytes_per_unit"] == 1024.0
    k = s["kernels"]["k_nn<4>"]
    assert k["fetch_bytes_per_launch"] == 250 * 2048  # the single-pair dispatch is dropped
    assert k["write_bytes_per_launch"] == 20 * 1024
    assert profiling.lookup(s, "k_nn") == 250 * 2048 + 20 * 1024
    assert profiling.lookup(s, "k_missing") is None
