"""Runs the C++ drop-in tests (query-engine_amd/host/tests/host_tests.cpp):
device operators behind the Operator contract and the plan builder."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "query-engine_amd" / "lib" / "qe_host_tests"


@pytest.mark.gpu
def test_host_operator_layer():
    assert EXE.exists(), "build with `python query-engine_amd/build.py`"
    # The executable is a build product (not tracked): refuse a stale one
    # rather than test old code.
    srcs = list((ROOT / "query-engine_amd" / "host").rglob("*.cpp")) + \
        list((ROOT / "query-engine_amd" / "host").glob("*.h")) + [ROOT / "include" / "qe_gpu.h"]
    newest = max(s.stat().st_mtime for s in srcs)
    assert EXE.stat().st_mtime >= newest, "qe_host_tests is older than its sources: rebuild"
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
