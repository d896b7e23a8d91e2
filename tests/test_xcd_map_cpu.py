"""Host check of the XCD-aware workgroup renumbering that every column kernel uses
(csrc/common.h xcd_spread / xcd_map): compiled with hipcc as host code and run on the CPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


def test_xcd_map_is_a_bijection(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "xcd_map_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", str(HERE / "xcd_map_check.cpp"), "-o", str(exe)], check=True,
                   capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() == "ok", out.stdout
