"""Trace, bench and profile output of real calls (layer mode 7) written to
ROCSOLVER_LOG_PATH: the bench line of every public call in the reference's
format, a trace tree per call with the stage the dispatcher chose, and
profile totals of the stages (library/src/include/rocsolver_logger.hpp:
160-300)."""
import numpy as np
import pytest

import rocsolver_amd as ra

pytestmark = pytest.mark.gpu


def test_trace_bench_profile(torch_gpu, handle, tmp_path, monkeypatch):
    torch = torch_gpu
    path = tmp_path / "rocsolver.log"
    monkeypatch.setenv("ROCSOLVER_LAYER", "7")
    monkeypatch.setenv("ROCSOLVER_LEVELS", "2")
    monkeypatch.setenv("ROCSOLVER_LOG_PATH", str(path))
    for k in ("ROCSOLVER_LOG_TRACE_PATH", "ROCSOLVER_LOG_BENCH_PATH", "ROCSOLVER_LOG_PROFILE_PATH"):
        monkeypatch.delenv(k, raising=False)
    assert ra.call("rocsolver_log_begin") == 0
    try:
        A = torch.randn((4, 32, 32), dtype=torch.float64, device="cuda")
        B = torch.randn((4, 2, 32), dtype=torch.float64, device="cuda")
        ipiv, info = ra.getrf(handle, A)
        ra.getrs(handle, A, ipiv, B)
        ra.getrf(handle, A)
        torch.cuda.synchronize()
    finally:
        assert ra.call("rocsolver_log_end") == 0
    text = path.read_text()
    assert ("./rocsolver-bench -f getrf_strided_batched -r d -m 32 -n 32 --lda 32 --strideA 1024 "
            "--strideP 32 --batch_count 4") in text
    assert ("./rocsolver-bench -f getrs_strided_batched -r d --trans N -n 32 --nrhs 2 --lda 32 "
            "--strideA 1024 --strideP 32 --ldb 32 --strideB 64 --batch_count 4") in text
    assert "------- ENTER rocsolver_dgetrf_strided_batched trace tree -------" in text
    assert "getrf_small (m: 32, n: 32, pivot: 1, bc: 4)" in text
    assert "------- EXIT rocsolver_dgetrs_strided_batched trace tree -------" in text
    prof = text.split("------- PROFILE -------", 1)[1]
    assert "getrf_small: Calls: 2, Total Time:" in prof
    assert "getrs_small: Calls: 1, Total Time:" in prof
