"""The request server's slot retirement (csrc/mqpiv_server_core.h), built for the
CPU by tests/server_core_test.cpp with injected event-query results: a slot
whose completion query fails (a sticky device error) ends its requests with
MQPIV_ERR_DEVICE, releases their pool memory, records the error and goes idle,
so that waiters and shutdown do not block; finished slots complete with
MQPIV_OK and running ones stay busy."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
GXX = shutil.which("g++")

MQPIV_OK, MQPIV_ERR_DEVICE = 0, 5


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    if not GXX:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("srv") / "server_core_test"
    r = subprocess.run([GXX, "-O1", "-std=c++17", "-Wall", "-Werror", str(HERE / "server_core_test.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    txt = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=60).stdout
    return {k: int(v) for k, v in (ln.split() for ln in txt.strip().splitlines())}


def test_failed_query_ends_requests_with_device_error(out):
    assert out["retired_any"] == 1
    assert out["status0"] == MQPIV_ERR_DEVICE and out["status1"] == MQPIV_ERR_DEVICE
    assert out["err_set"] == 1
    assert out["busy0"] == 0


def test_done_and_running_slots(out):
    assert out["status2"] == MQPIV_OK and out["busy1"] == 0
    assert out["status3"] == -1 and out["busy2"] == 1
    assert out["polls"] == 3
    # only the successful request counts toward throughput
    assert out["pairs_done"] == 1 and out["ia_done"] == 300


def test_pool_memory_is_released(out):
    assert out["used_before"] == 4096 * (1 + 2 + 3 + 4)
    assert out["used_after"] == 4096 * 4  # the running request keeps its block
    assert out["status3_later"] == MQPIV_ERR_DEVICE and out["busy2_later"] == 0
    assert out["used_final"] == 0 and out["realloc_whole"] == 1
    assert out["polls_idle"] == 0
