"""Device code generation guard (CPU: cross-compiles for gfx950).

This ROCm toolchain emits ``s_mov_b64 sN, 0x7ff0000000000000`` (a uniform
FP64 constant with a zero low half, e.g. +inf) as a single 32-bit literal
holding the low half, so the scalar registers read 0.0 instead of +inf.  The
kNN grid kernel once lost its empty-cell test to it.  Every library source is
compiled to gfx950 assembly with the build's flags, and no scalar instruction
may carry a literal wider than 32 bits."""
import concurrent.futures as cf
import re
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "mpt_amd"))
import build  # noqa: E402

WIDE_SCALAR_LITERAL = re.compile(r"^\s+s_[a-z0-9_]+\s.*\b0x[0-9a-fA-F]{9,}\b", re.M)


@pytest.mark.skipif(shutil.which(build.HIPCC) is None and not Path(build.HIPCC).exists(), reason="no hipcc")
def test_no_wide_scalar_literals(tmp_path):
    flags = [f for f in build.HIP_FLAGS if f != "-fPIC"]

    def asm(src):
        out = tmp_path / (src + ".s")
        subprocess.run([build.HIPCC, *flags, "--cuda-device-only", "-S", str(build.CSRC / src), "-o", str(out)],
                       check=True, capture_output=True)
        return src, WIDE_SCALAR_LITERAL.findall(out.read_text())

    with cf.ThreadPoolExecutor(8) as ex:
        bad = {src: hits[:3] for src, hits in ex.map(asm, build.LIB_SOURCES) if hits}
    assert not bad, bad
