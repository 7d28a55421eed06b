"""The C++ host layer on the CPU: binary splitting, decimal conversion, pi
digits and dispatch rules (tests/cpp/test_host.cpp), and the pi program's
GMP-only digit files against the reference program built from its sources."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "pi_compute_amd" / "bin"
GOLDEN = ROOT / "tests" / "golden" / "pi_1000.txt"
REF_PI = ROOT / "oracle" / "_ref" / "pi_compute_ref"


def test_host_layer():
    exe = BIN / "test_host"
    assert exe.exists(), "host test not built (run __graft_entry__.build())"
    r = subprocess.run([str(exe), str(GOLDEN)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0 and "0 failure(s)" in r.stdout, r.stdout + r.stderr


def _pi(tmp_path, digits, *args):
    out = tmp_path / f"pi_{digits}_{'_'.join(a.strip('-') for a in args) or 'default'}.txt"
    r = subprocess.run([str(BIN / "pi_compute_hip"), "--digits", str(digits), "--output", str(out), *args],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return out.read_bytes()


def test_pi_program_file_format(tmp_path):
    got = _pi(tmp_path, 1000)
    assert got == GOLDEN.read_bytes()  # "3." + 1000 digits, no newline
    assert _pi(tmp_path, 1) == b"3.1"
    assert _pi(tmp_path, 7, "--tree", "recursive") == b"3.1415926"


@pytest.mark.parametrize("digits", [54321, 300000])
def test_pi_program_matches_reference_program(tmp_path, digits):
    if not REF_PI.exists():
        pytest.skip("reference program not built here (needs /root/reference)")
    ref = tmp_path / "ref.txt"
    r = subprocess.run([str(REF_PI), "--digits", str(digits), "--output", str(ref)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    want = ref.read_bytes()
    assert _pi(tmp_path, digits) == want
    assert _pi(tmp_path, digits, "--tree", "recursive") == want
