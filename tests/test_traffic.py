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
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Value"])
        w.writeheader()
        for name, kib in rows:
            w.writerow({"Kernel_Name": name, "Counter_Value": kib})


def test_parse_scales_by_the_copy(tmp_path):
    # counters in KiB: the copy reads 1 GiB but FETCH_SIZE reports half of it
    # (scale 2); WRITE_SIZE reports it whole (scale 1)
    gib_kib = (1 << 30) / 1024
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(fetch, [(COPY, gib_kib / 2), (LU, 1000.0), (LU, 3000.0), (OTHER, 5.0)])
    _write(write, [(COPY, gib_kib), (LU, 4000.0), (LU, 4000.0), (OTHER, 5.0)])
    res = traffic.parse(str(fetch), str(write))
    assert res["copy_kernel"] == COPY
    assert abs(res["fetch_scale"] - 2.0) < 1e-12 and abs(res["write_scale"] - 1.0) < 1e-12
    assert list(res["kernels"]) == [LU]  # only this library's kernels are reported
    k = res["kernels"][LU]
    assert k["dispatches"] == 2
    assert k["fetch_bytes"] == 2000.0 * 1024 * 2  # mean per dispatch, KiB -> bytes, scaled
    assert k["write_bytes"] == 4000.0 * 1024


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
