"""C ABI: the library loads, exports every entry point declared in
include/mptg.h, and reports errors without a GPU (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import numpy as np
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


def test_create_reports_errors_through_the_abi():
    import mpt_amd
    try:
        ctx = mpt_amd.Context(0)
    except mpt_amd.MptgError as e:  # no device here: a clean error, not a crash
        assert "device" in str(e).lower() or "hip" in str(e).lower()
        return
    with pytest.raises(mpt_amd.MptgError):  # with a device: an out-of-order call is rejected
        ctx.knn(np.zeros((1, 2)), 1)
    ctx.close()
