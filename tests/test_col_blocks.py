"""Workgroup order of the forward column kernels with tiles of 2^5 columns or
more (prime_grouped_block in pi_compute_amd/csrc/ntt_device.h).

The mapping is read from the C++ source: tests/cpp/col_blocks.cpp is built
host-only and prints the (tile, prime, job) of every linear workgroup.  Checked
here: it is a bijection onto the grid for every shape (ragged tails included),
and wherever a whole group of 8 (tile, job) pairs exists, the three primes of
a pair run on one XCD (linear index mod 8) within 24 consecutive workgroups.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "pi_compute_amd" / "csrc"

SHAPES = [(1, 1), (1, 7), (2, 3), (4, 1), (4, 3071), (4, 5), (3, 11), (8, 36), (16, 9), (32, 587),
          (5, 13), (7, 8), (1024, 3), (2, 65535)]


@pytest.fixture(scope="module")
def grids(tmp_path_factory):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if Path("/opt/rocm/bin/hipcc").exists() else None)
    if hipcc is None:
        pytest.skip("hipcc is not installed")
    exe = tmp_path_factory.mktemp("colblk") / "col_blocks"
    subprocess.run([hipcc, "--offload-host-only", "-std=c++17", "-I", str(CSRC),
                    str(ROOT / "tests" / "cpp" / "col_blocks.cpp"), "-o", str(exe)], check=True)
    args = [str(v) for nx, nz in SHAPES for v in (nx, nz)]
    out = subprocess.run([str(exe)] + args, check=True, capture_output=True, text=True).stdout
    blocks = out.split("\n\n")
    res = {}
    for shp, text in zip(SHAPES, blocks):
        res[shp] = [tuple(int(v) for v in line.split()) for line in text.strip().split("\n") if line]
    return res


@pytest.mark.parametrize("shape", SHAPES)
def test_bijection(grids, shape):
    nx, nz = shape
    got = grids[shape]
    assert len(got) == 3 * nx * nz
    assert all(0 <= t < nx and 0 <= p < 3 and 0 <= j < nz for t, p, j in got)
    assert len(set(got)) == len(got)


@pytest.mark.parametrize("shape", SHAPES)
def test_primes_of_a_pair_share_an_xcd(grids, shape):
    nx, nz = shape
    got = grids[shape]
    where = {blk: L for L, blk in enumerate(got)}
    full = (nx * nz) // 8
    for u in range(8 * full):
        t, j = u % nx, u // nx
        ls = [where[(t, p, j)] for p in range(3)]
        assert len({L % 8 for L in ls}) == 1, (u, ls)
        assert max(ls) - min(ls) < 24, (u, ls)
    # The tail keeps the primes of a pair next to each other.
    for u in range(8 * full, nx * nz):
        t, j = u % nx, u // nx
        ls = sorted(where[(t, p, j)] for p in range(3))
        assert ls[2] - ls[0] == 2


def test_planted_mistake_is_caught():
    """A mapping with the prime taken from L mod 3 instead of the XCD slot spreads
    the primes of a pair over XCDs: the property check must reject it."""
    nx, nz = 4, 8
    got = []
    for L in range(3 * nx * nz):
        u, prime = L // 3, L % 3
        got.append((u % nx, prime, u // nx))
    where = {blk: L for L, blk in enumerate(got)}
    xcds = {where[(0, p, 0)] % 8 for p in range(3)}
    assert len(xcds) > 1
