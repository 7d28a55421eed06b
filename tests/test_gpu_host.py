"""The C++ drop-in (HipNttMultiplier) against GmpMultiplier, through the
compiled host test (tests/cpp/test_hip_multiplier.cpp)."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
pytestmark = pytest.mark.gpu


def test_hip_multiplier_matches_gmp():
    exe = ROOT / "pi_compute_amd" / "bin" / "test_hip_multiplier"
    assert exe.exists(), "host test not built (run __graft_entry__.build())"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
