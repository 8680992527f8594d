"""Bisection-tree index helpers of the surrogate order (CPU: a host program
built with hipcc against mpt_amd/csrc/bisect_dev.hpp; see
tests/bisect_index_check.cpp for what it checks)."""
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "mpt_amd"))
import build  # noqa: E402


@pytest.mark.skipif(shutil.which(build.HIPCC) is None and not Path(build.HIPCC).exists(), reason="no hipcc")
def test_bisect_index_helpers(tmp_path):
    exe = tmp_path / "bisect_index_check"
    subprocess.run([build.HIPCC, "-std=c++20", "-O1", "-I", str(build.CSRC), "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "bisect_index_check.cpp"), "-o", str(exe)], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "bad=0", r.stdout + r.stderr
