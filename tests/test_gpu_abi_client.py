"""The C ABI from a plain C++ process (no Python, no torch): the client built next to
libsav_hip.so runs a single-vesicle simulation and a two-member ensemble and checks them
(exit status 0)."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
CLIENT = Path(__file__).resolve().parents[1] / "julia-vesicles-growth-and-shrinkage_amd" / "sav_abi_client"


def test_cpp_abi_client():
    if not CLIENT.exists():
        raise FileNotFoundError(f"{CLIENT} is not built (run __graft_entry__.build())")
    r = subprocess.run([str(CLIENT)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "abi client ok" in r.stdout
