"""The library's host code (YAML reader, sensor factory, scene / emitter /
directivity builders, ping-file writer and readers, pose CSV reader) and the
CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer, leak checks
on: tests/sanitize builds them with instrumented host code and runs
host_checks.cpp over well-formed and malformed inputs (no GPU)."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "sanitize"


@pytest.mark.skipif(not Path("/opt/rocm/bin/hipcc").exists() or shutil.which("make") is None,
                    reason="needs hipcc and make")
def test_host_code_under_asan_and_ubsan():
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    r = subprocess.run(["make", "-C", str(HERE), f"-j{jobs}", "run"], capture_output=True, text=True,
                       timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "host checks ok" in r.stdout, tail
