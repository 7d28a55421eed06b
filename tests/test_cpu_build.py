"""The CPU-only build (hosts without ROCm) and the sanitizer builds of the host
layer, on the CPU.

* `make -C pi_compute_amd cpu` builds the pi program and the host test with
  g++ alone (hipcc replaced by /bin/false proves it is never invoked); the
  programs link no HIP runtime and no libpintt.so; GMP-only digits are the
  reference's, and --gpu fails at start-up with a no-device error, as the
  reference's GPU multipliers do without a device
  (src/arithmetic/int_ntt_multiplier.cpp:26-38; gating at
  src/arithmetic/gpu_ntt_multiplier.h:14).
* `make -C pi_compute_amd sanitize` builds the host test under
  AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer;
  both must pass with no report (binary splitting, decimal conversion and the
  final step run on host threads, src/engine/binary_splitting.cpp:183-193).
"""
import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "pi_compute_amd"
CPU = PKG / "cpu_bin"
GOLDEN = ROOT / "tests" / "golden" / "pi_1000.txt"


def _make(target):
    jobs = str(min(8, os.cpu_count() or 2))
    r = subprocess.run(["make", "-C", str(PKG), "-j", jobs, target, "HIPCC=/bin/false"], capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.fixture(scope="module")
def cpu_build():
    _make("cpu")
    return CPU


def test_cpu_only_programs_link_no_gpu_code(cpu_build):
    for exe in ("pi_compute_hip", "test_host"):
        r = subprocess.run(["ldd", str(cpu_build / exe)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0
        assert "libpintt" not in r.stdout and "amdhip" not in r.stdout, r.stdout


def test_cpu_only_pi_program(cpu_build, tmp_path):
    out = tmp_path / "pi.txt"
    r = subprocess.run([str(cpu_build / "pi_compute_hip"), "--digits", "1000", "--output", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == GOLDEN.read_bytes()
    r = subprocess.run([str(cpu_build / "pi_compute_hip"), "--digits", "1000", "--gpu", "--output", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "CPU-only build" in r.stderr, r.stdout + r.stderr


def test_cpu_only_host_tests(cpu_build):
    r = subprocess.run([str(cpu_build / "test_host"), str(GOLDEN)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "0 failure(s)" in r.stdout, r.stdout + r.stderr


def test_no_device_object_defines_the_whole_abi(cpu_build):
    from test_abi import _declared_symbols

    r = subprocess.run(["nm", "-g", "--defined-only", str(PKG / "_build" / "host" / "no_device.o")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    defined = set(re.findall(r"\bT (gpu_\w+)", r.stdout))
    assert defined == _declared_symbols()


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_layer_under_sanitizers(kind):
    _make("sanitize")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(CPU / f"test_host_{kind}"), str(GOLDEN)], capture_output=True, text=True, timeout=900,
                       env=env)
    report = r.stderr
    assert r.returncode == 0 and "0 failure(s)" in r.stdout, r.stdout[-3000:] + report[-3000:]
    assert "Sanitizer" not in report and "runtime error" not in report, report[-3000:]
