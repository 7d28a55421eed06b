"""Pi digit files through the GPU multiplier equal the GMP-only run, byte for
byte, in every dispatch mode (level batches, per-call, depth-first), with the
Newton final step and the Newton-quotient decimal conversion (the defaults
with --gpu; with device-resident levels the final step and the conversion
run on device numbers too, or on the host with --host-final-step and
--host-conversion-levels) and with the reference's mpf final step."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "pi_compute_amd" / "bin"
REF_PI = ROOT / "oracle" / "_ref" / "pi_compute_ref"
MODEL = ROOT / "pi_compute_amd" / "data" / "cost_model.txt"
pytestmark = pytest.mark.gpu


def _run(args, out):
    print("run:", " ".join(args), flush=True)
    r = subprocess.run([*args, "--output", str(out)], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip().splitlines()[-2], flush=True)
    return out.read_bytes()


@pytest.fixture(scope="module")
def gmp_digits(tmp_path_factory):
    d = tmp_path_factory.mktemp("pi")
    out = {}
    for digits in (1000000,):
        exe = [str(REF_PI)] if REF_PI.exists() else [str(BIN / "pi_compute_hip")]
        out[digits] = _run([*exe, "--digits", str(digits)], d / f"gmp_{digits}.txt")
    return out


@pytest.mark.parametrize(
    "mode",
    [
        ["--tree", "levels"],
        ["--tree", "levels", "--per-call"],
        ["--tree", "recursive"],
        ["--tree", "levels", "--gpu-threshold", "64"],
        ["--tree", "recursive", "--gpu-threshold", "64"],
        ["--tree", "levels", "--final", "mpf", "--convert", "host"],
        ["--tree", "levels", "--gpu-threshold", "64", "--convert", "multiplier"],
        ["--tree", "levels", "--device-levels"],
        ["--tree", "levels", "--device-levels", "2000"],
        ["--tree", "levels", "--device-levels", "--bs-segments", "3"],
        ["--tree", "levels", "--device-levels", "--bs-segments", "1"],
        ["--tree", "levels", "--device-levels", "--no-overlap"],
        ["--tree", "levels", "--device-levels", "--host-conversion-levels"],
        ["--tree", "levels", "--device-levels", "--host-final-step"],
        ["--tree", "levels", "--device-levels", "--convert-leaf", "4096"],
    ],
    ids=["levels", "levels-per-call", "recursive", "levels-t64", "recursive-t64", "levels-mpf-host",
         "levels-t64-convert", "device-levels", "device-levels-2000", "device-levels-3-segments",
         "device-levels-one-tree", "device-levels-serial-constants", "device-levels-host-conversion",
         "device-levels-host-final-step", "device-levels-conversion-leaf-4096"],
)
def test_pi_gpu_matches_gmp_run(tmp_path, gmp_digits, mode):
    digits = 1000000
    got = _run([str(BIN / "pi_compute_hip"), "--digits", str(digits), "--gpu", *mode], tmp_path / "gpu.txt")
    assert got == gmp_digits[digits]


@pytest.mark.parametrize("device", [False, True], ids=["host-levels", "device-levels-auto"])
def test_pi_gpu_learned_dispatch_matches_gmp_run(tmp_path, gmp_digits, device):
    if not MODEL.exists():
        pytest.skip("no trained cost model in the tree")
    extra = ["--device-levels", "auto"] if device else []
    got = _run([str(BIN / "pi_compute_hip"), "--digits", "1000000", "--gpu", "--dispatch", "mlp", "--model",
                str(MODEL), *extra], tmp_path / "mlp.txt")
    assert got == gmp_digits[1000000]


@pytest.mark.parametrize("digits", [8193, 65537, 16777216])
def test_pi_device_tail_at_power_of_two_widths(tmp_path, digits):
    """The device-resident tail (final step, conversion levels and leaves)
    where the digit count of pi * 10^D sits just past a power of two (the
    number of conversion levels changes there): the reference's own GMP
    program gives the expected file."""
    if not REF_PI.exists():
        pytest.skip("the reference program is not built in this tree")
    want = _run([str(REF_PI), "--digits", str(digits)], tmp_path / "ref.txt")
    got = _run([str(BIN / "pi_compute_hip"), "--digits", str(digits), "--gpu", "--tree", "levels", "--device-levels"],
               tmp_path / "gpu.txt")
    assert got == want
