"""The C-ABI library loads without a GPU and exports every function
include/qe_gpu.h declares (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"\b(qe_gpu_[a-z0-9_]+)\s*\(", h.read_text()))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "qe_gpu_join_aggregate" in names and "qe_gpu_asof_probe" in names
    assert len(names) >= 50


def test_library_exports_every_declared_symbol(qe):
    L = ctypes.CDLL(str(qe.LIB_PATH))
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol(qe):
    assert set(declared()) <= set(qe._SIGS), sorted(set(declared()) - set(qe._SIGS))


def test_status_codes_map_to_reference_exception_types(qe):
    assert issubclass(qe.InvalidArgument, ValueError)  # std::invalid_argument
    assert issubclass(qe.Overflow, OverflowError)       # std::overflow_error
    assert issubclass(qe.OutOfMemory, MemoryError)      # std::bad_alloc


def test_context_without_gpu_fails_loudly(qe):
    import os
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    with pytest.raises(qe.QeError):
        qe.Context(0)
