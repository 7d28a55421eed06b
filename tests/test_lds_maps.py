"""LDS bank maps of the radix-16 tile kernels (pi_compute_amd/csrc/ntt_device.h
lds_map, Tile::word, Tile::word_at; used by k_col_fwd, k_col_inv and
k_row_conv).

The maps are read from the C++ source itself: tests/cpp/lds_maps.cpp is built
host-only from ntt_device.h and prints lds_map for every tile shape the
library compiles (parsed from the PINTT_*_CASES lists).  On a lane-by-lane
model of every LDS access those kernels make -- the 16 loads/stores of a
thread at positions e NT + tid, and the 2^K accesses base + i DD of each group
of each radix-16 stage (Stage::locate) -- each map must:

* be a permutation of the tile's words (no word shared, none outside WORDS);
* give every access the word Tile::word_at computes from the group's first
  word (the kernels never evaluate word() on the sum, so a map that is not
  linear over GF(2), or an access whose offset shares bits with its base,
  would read the wrong word);
* put the 32 lanes of each half-wave on 32 distinct banks (32 banks of 4
  bytes, word mod 32; the hardware counter SQ_LDS_BANK_CONFLICT reads zero
  for these kernels).

The planted cases show the checks bite: the identity layout has conflicts,
and a padded layout (x + x / 16) breaks the word_at decomposition.
"""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "pi_compute_amd" / "csrc"
KE = 16  # elements per thread


def _cases(src, macro):
    text = (CSRC / src).read_text()
    m = re.search(r"#define " + macro + r"\(X\)((?:[^\n]*\\\n)*[^\n]*)", text)
    assert m, macro
    return {(int(a), int(b)) for a, b in re.findall(r"X\((\d+),\s*(\d+)\)", m.group(1))}


def _shapes():
    cols = _cases("ntt_cols.hip", "PINTT_COL_CASES_NARROW") | _cases("ntt_cols.hip", "PINTT_COL_CASES_WIDE")
    rows = (_cases("ntt_rows.hip", "PINTT_ROW_CASES_NARROW") | _cases("ntt_rows.hip", "PINTT_ROW_CASES_WIDE")
            | _cases("ntt_rows.hip", "PINTT_PAIR_CASES_NARROW"))
    return sorted((l, s, True) for l, s in cols) + sorted((l, s, False) for l, s in rows)


@pytest.fixture(scope="module")
def maps(tmp_path_factory):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if Path("/opt/rocm/bin/hipcc").exists() else None)
    if hipcc is None:
        pytest.skip("hipcc is not installed")
    exe = tmp_path_factory.mktemp("lds") / "lds_maps"
    subprocess.run([hipcc, "--offload-host-only", "-std=c++17", "-I", str(CSRC),
                    str(ROOT / "tests" / "cpp" / "lds_maps.cpp"), "-o", str(exe)], check=True)
    shapes = _shapes()
    args = [str(v) for shp in shapes for v in (shp[0], shp[1], int(shp[2]))]
    out = subprocess.run([str(exe)] + args, check=True, capture_output=True, text=True).stdout.split("\n")
    return {shp: tuple(int(v) for v in line.split()) for shp, line in zip(shapes, out)}


def xor_map(terms):
    s1, m1, t1, s2, m2, t2 = terms

    def word(x):
        x = np.asarray(x, dtype=np.int64)
        return x ^ (((x >> s1) & m1) << t1) ^ (((x >> s2) & m2) << t2)
    return word


def word_at(word, w0, c):
    wc = int(word(c))
    return (w0 ^ (wc & 31)) + (wc & ~31)


def accesses(logl, logseq, cols):
    """(x0, c) of every LDS instruction: lane t touches position x0[t] + c."""
    L, SEQ = 1 << logl, 1 << logseq
    NT = L * SEQ // KE
    t = np.arange(NT, dtype=np.int64)
    out = [(t, e * NT) for e in range(KE)]
    s0 = 0
    while s0 < logl:
        K = min(logl - s0, 4)
        G, DD = 1 << K, L >> (s0 + K)
        GPS = L // G
        for q in range(KE // G):
            gi = q * NT + t
            if cols:
                seq, grp = gi & (SEQ - 1), gi >> logseq
            else:
                grp, seq = gi & (GPS - 1), gi // GPS
            base = (grp // DD) * (DD * G) + grp % DD
            x0 = (base << logseq) + seq if cols else (seq << logl) + base
            stride = SEQ if cols else 1
            out += [(x0, i * DD * stride) for i in range(G)]
        s0 += K
    return out


def conflict_cycles(words):
    """Extra LDS cycles of one instruction: per half-wave, the most lanes on one bank, minus one."""
    cyc = 0
    for h in range(0, len(words), 32):
        cnt = np.bincount(words[h:h + 32] % 32, minlength=32)
        cyc += int(cnt.max()) - 1
    return cyc


def check(shape, word, decompose=True):
    logl, logseq, cols = shape
    n = 1 << (logl + logseq)
    img = word(np.arange(n))
    assert len(np.unique(img)) == n and img.max() < n, "not a permutation of the tile"
    cyc = 0
    for x0, c in accesses(logl, logseq, cols):
        assert not np.any(x0 & c), "offset shares bits with the base"
        w = word(x0 + c)
        if decompose:
            assert np.array_equal(word_at(word, word(x0), c), w), "word_at differs from word"
        cyc += conflict_cycles(w)
    return cyc


def test_every_compiled_shape_has_a_map(maps):
    assert len(maps) == len(_shapes()) >= 40
    for terms in maps.values():
        s1, m1, t1, s2, m2, t2 = terms
        assert (m1 << t1) < 32 and (m2 << t2) < 32  # only bits 0..4 change


def test_maps_are_conflict_free_permutations(maps):
    for shape, terms in maps.items():
        assert check(shape, xor_map(terms)) == 0, shape


def test_identity_layout_conflicts():
    # The planted layouts: the identity puts radix-16 strides on few banks.
    assert check((7, 3, True), lambda x: np.asarray(x, dtype=np.int64)) > 0
    assert check((8, 3, False), lambda x: np.asarray(x, dtype=np.int64)) > 0


def test_padded_layout_breaks_word_at():
    def pad(x):
        x = np.asarray(x, dtype=np.int64)
        return x + (x >> 4)
    with pytest.raises(AssertionError, match="not a permutation"):
        check((8, 3, False), pad)  # needs WORDS + WORDS / 16 words
    # with room for the pad words, the decomposition is what fails
    bad = 0
    for x0, c in accesses(8, 3, False):
        bad += int(np.sum(word_at(pad, pad(x0), c) != pad(x0 + c)))
    assert bad > 0
