"""Host logic behind the bench line's roofline.traffic: the counter parser's
calibration against the 1 GiB copy, and the committed counter summaries that
bench.py reads (no GPU)."""
import csv
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from rocsolver_amd import traffic  # noqa: E402

COPY = ("void at::native::vectorized_elementwise_kernel<4, at::native::AUnaryFunctor<double, double, "
        "double, at::native::binary_internal::MulFunctor<double> > >(int)")
LU = "void rsa::getrf_small_kernel<double, 32, 32, true, true>(int)"
OTHER = "void at::native::fill_kernel(int)"


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Grid_Size", "Counter_Value"])
        w.writeheader()
        for name, grid, kib in rows:
            w.writerow({"Kernel_Name": name, "Grid_Size": grid, "Counter_Value": kib})


def test_parse_scales_by_the_copy(tmp_path):
    # counters in KiB: the copy reads 1 GiB but FETCH_SIZE reports half of it
    # (scale 2); WRITE_SIZE reports it whole (scale 1)
    gib_kib = (1 << 30) / 1024
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(fetch, [(COPY, 1 << 27, gib_kib / 2), (LU, 4096, 1000.0), (LU, 4096, 3000.0),
                   (LU, 2048, 700.0), (OTHER, 64, 5.0)])
    _write(write, [(COPY, 1 << 27, gib_kib), (LU, 4096, 4000.0), (LU, 4096, 4000.0),
                   (LU, 2048, 900.0), (OTHER, 64, 5.0)])
    res = traffic.parse(str(fetch), str(write))
    assert res["copy_kernel"] == COPY
    assert abs(res["fetch_scale"] - 2.0) < 1e-12 and abs(res["write_scale"] - 1.0) < 1e-12
    # only this library's kernels, one entry per grid size (a launch over part
    # of the batch is not averaged with one over all of it)
    ks = res["kernels"]
    assert [(k["kernel"], k["grid"]) for k in ks] == [
        ("void rsa::getrf_small_kernel<double, 32, 32, true, true>", 2048),
        ("void rsa::getrf_small_kernel<double, 32, 32, true, true>", 4096)]
    assert ks[0]["dispatches"] == 1 and ks[1]["dispatches"] == 2
    assert ks[1]["fetch_bytes"] == 2000.0 * 1024 * 2  # mean per dispatch, KiB -> bytes, scaled
    assert ks[1]["write_bytes"] == 4000.0 * 1024
    assert ks[0]["fetch_bytes"] == 700.0 * 1024 * 2 and ks[0]["write_bytes"] == 900.0 * 1024


def test_committed_summaries_are_read():
    mixed = bench.measured_traffic("lu_fused_kernel<double,float,256>")
    n32 = bench.measured_traffic("getrf_small_kernel<double,32,32,pivot>",
                                 "traffic_getrf_n32_batch131072.md")
    assert mixed is not None and n32 is not None
    # the n=32 factorisation moves its algorithmic bytes (A read and written,
    # ipiv, info) to within 1 %
    algorithmic = 131072 * (2 * 32 * 32 * 8 + 32 * 4 + 4)
    assert abs(n32 - algorithmic) < 0.01 * algorithmic
    # the FP32 LU of the n=256 batch reads at least its FP64 input and writes
    # at least its FP32 factors
    assert mixed >= 4096 * 256 * 256 * (8 + 4)


def test_missing_summary_gives_none():
    assert bench.measured_traffic("getrf_small_kernel<double,32,32,pivot>", "no_such_file.md") is None
    assert bench.measured_traffic("no_such_kernel") is None
