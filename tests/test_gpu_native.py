"""The batched, served and tuned forms driven from C++ through the C ABI alone
(tests/native/abi_drive.cpp): ihe_batch_* (eager and graph replays), the
ihe_server_* request loop (one request in two chunks) and ihe_band_create_tuned,
each bit-exact against the single-cohort entry points, which the other GPU
tests pin to the oracle. IHE_NATIVE_DIR=asan runs the AddressSanitizer/UBSan
build instead (make -C tests/native asan)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batch_server_and_tuner_from_cpp():
    exe = os.path.join(ROOT, "tests", "native", os.environ.get("IHE_NATIVE_DIR", ""), "abi_drive")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
    lines = r.stdout.splitlines()
    assert sum(line.startswith("ok ") for line in lines) == 8, r.stdout
    assert not any(line.startswith("FAIL") for line in lines), r.stdout
