"""The logging session API (rocsolver_log_*), host-only parts: status codes
of the session calls and the environment variables read by log_begin, as the
reference logger defines them (library/src/common/rocsolver_logger.cpp:
191-330).  The trace/bench/profile output of real calls is checked on the
GPU (test_gpu_logging.py)."""
import os

import pytest

import rocsolver_amd as ra

INTERNAL_ERROR, INVALID_VALUE = 6, 11


@pytest.fixture
def clean_env(monkeypatch):
    for k in ("ROCSOLVER_LAYER", "ROCSOLVER_LEVELS", "ROCSOLVER_LOG_PATH",
              "ROCSOLVER_LOG_TRACE_PATH", "ROCSOLVER_LOG_BENCH_PATH", "ROCSOLVER_LOG_PROFILE_PATH"):
        monkeypatch.delenv(k, raising=False)
    yield monkeypatch
    ra.call("rocsolver_log_end")  # leave no session behind


def test_session_calls_need_an_active_session(clean_env):
    for f in ("rocsolver_log_end", "rocsolver_log_restore_defaults",
              "rocsolver_log_write_profile", "rocsolver_log_flush_profile"):
        assert ra.call(f) == INTERNAL_ERROR
    assert ra.call("rocsolver_log_set_layer_mode", 7) == INTERNAL_ERROR
    assert ra.call("rocsolver_log_set_max_levels", 2) == INTERNAL_ERROR


def test_session_lifecycle(clean_env):
    assert ra.call("rocsolver_log_begin") == 0
    assert ra.call("rocsolver_log_begin") == INTERNAL_ERROR  # one session at a time
    assert ra.call("rocsolver_log_set_max_levels", 0) == INVALID_VALUE
    assert ra.call("rocsolver_log_set_max_levels", 3) == 0
    assert ra.call("rocsolver_log_set_layer_mode", 7) == 0
    assert ra.call("rocsolver_log_write_profile") == 0
    assert ra.call("rocsolver_log_flush_profile") == 0
    assert ra.call("rocsolver_log_restore_defaults") == 0
    assert ra.call("rocsolver_log_end") == 0
    assert ra.call("rocsolver_log_end") == INTERNAL_ERROR


def test_environment_variables(clean_env, tmp_path):
    path = tmp_path / "log.txt"
    clean_env.setenv("ROCSOLVER_LAYER", "0x7")
    clean_env.setenv("ROCSOLVER_LEVELS", "2")
    clean_env.setenv("ROCSOLVER_LOG_PATH", str(path))
    assert ra.call("rocsolver_log_begin") == 0
    assert ra.call("rocsolver_log_end") == 0
    text = path.read_text()
    assert text.startswith("ROCSOLVER LOG FILE\nrocSOLVER Version: ")  # header once per file
    assert text.count("ROCSOLVER LOG FILE") == 1
    # out-of-range values make log_begin fail (the session still has to be ended)
    clean_env.setenv("ROCSOLVER_LAYER", "-1")
    assert ra.call("rocsolver_log_begin") == INTERNAL_ERROR
    assert ra.call("rocsolver_log_end") == 0
    clean_env.setenv("ROCSOLVER_LAYER", "1")
    clean_env.setenv("ROCSOLVER_LEVELS", "0")
    assert ra.call("rocsolver_log_begin") == INTERNAL_ERROR
    assert ra.call("rocsolver_log_end") == 0
