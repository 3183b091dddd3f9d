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
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
