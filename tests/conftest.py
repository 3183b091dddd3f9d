import importlib.util
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "tests"))

# The CI seed of the reference suites (CMakeLists.txt:461-462); override with
# QE_SEED=<n> to replay a randomized case.
SEED = int(os.environ.get("QE_SEED", "20260614"))


def load_package():
    """Import query-engine_amd/ (a directory name that is not an identifier)."""
    name = "query_engine_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, ROOT / "query-engine_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    print(f"[qe] seed={SEED} (replay: QE_SEED={SEED})")


@pytest.fixture(scope="session")
def qe():
    return load_package()


@pytest.fixture(scope="session")
def ctx(qe):
    # With a GPU present the HIP library must load and a context must open;
    # any failure here fails the GPU tests loudly (there is no CPU fallback).
    c = qe.Context(0)
    yield c
    c.close()


@pytest.fixture
def rng():
    return np.random.default_rng(SEED)
