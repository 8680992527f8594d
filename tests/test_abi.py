"""C ABI: the library loads, exports every entry point declared in
include/mptg.h, and reports errors without a GPU (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"\b(mptg_[a-z0-9_]+)\s*\(", h.read_text()))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "mptg_knn" in names and "mptg_link" in names and "mptg_prrt_run" in names


def test_library_exports_every_declared_symbol():
    from mpt_amd import _native
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from mpt_amd import _native
    assert sorted(_native.EXPORTED) == _declared()


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import mpt_amd
    with pytest.raises(mpt_amd.MptgError):
        mpt_amd.Context(0)
