"""The carry scheme of the CRT + carry pass (ntt_carry.hip: k_crt_carry,
k_carry_fixup), restated in Python over word sums and checked against a plain
sequential carry (the reference's host loop, src/gpu/int_ntt.cu:352-377).

Word sums s_j < 6 * 2^32 (at most two chunk products of three CRT words).  A
run of words maps a carry-in c in {0..7} to b + [c >= t] (t = 8: never).  A
workgroup sums kCarryWindowWords (32) words before its block and the block's
2016 words, eight per thread; when the window's composed map is constant it
gives the block's carry-in exactly, else the block is written with carry-in 0,
flagged, and the fixup composes the earlier blocks' maps and ripples the true
carry-in through the block's all-ones words.  Random sums, sums whose low
halves are all ones (every window unresolved), and mixtures are all checked;
planted mistakes in the map rules fail.
"""
import random

import pytest

M = 1 << 32
WT, THREADS, WINDOW = 8, 256, 32
WB = THREADS * WT - WINDOW  # words a block writes
IDENT = None


def make(b, t):
    return (b, t)


def apply(f, c):
    return c if f is IDENT else f[0] + (1 if c >= f[1] else 0)


def compose(later, earlier):
    if earlier is IDENT:
        return later
    if later is IDENT:
        return earlier
    (gb, gt), (fb, ft) = earlier, later
    lo = fb + (1 if gb >= ft else 0)
    hi = fb + (1 if gb + 1 >= ft else 0)
    return make(lo, 8 if (gt >= 8 or lo == hi) else gt)


def thread_map(sums):
    """The kernel's map of eight consecutive word sums (carry-in 0 chain)."""
    c, w0, ones = 0, 0, True
    for e, s in enumerate(sums):
        v = s + c
        o, c = v % M, v >> 32
        if e == 0:
            w0 = o
        else:
            ones = ones and o == M - 1
    room = (M - w0) % M
    return make(c, room if (ones and w0 != 0 and room <= 7) else 8)


def sequential(sums):
    out, c = [], 0
    for s in sums:
        v = s + c
        out.append(v % M)
        c = v >> 32
    assert c == 0
    return out


def blocked(sums, mistake=None):
    """The two kernels over the sums (len(sums) words, the product's size)."""
    n = len(sums)
    nblk = -(-n // WB)
    out = [None] * n
    maps, flagged = [], []
    word = lambda j: sums[j] if 0 <= j < n else 0  # noqa: E731
    for lb in range(nblk):
        first = lb * WB - WINDOW
        fs = [thread_map([word(first + WT * t + e) for e in range(WT)]) for t in range(THREADS)]
        win = IDENT
        for t in range(WINDOW // WT):
            win = compose(fs[t], win)
        resolved = win[1] >= 8
        if mistake == "window_always_resolved":
            resolved = True
        cin = win[0] if resolved else 0
        acc = IDENT
        for t in range(WINDOW // WT, THREADS):
            c = apply(acc, cin)
            for e in range(WT):
                j = first + WT * t + e
                v = word(j) + c
                if j < n:
                    out[j] = v % M
                c = v >> 32
            acc = compose(fs[t], acc)
        maps.append(acc)
        flagged.append(not resolved)
    # k_carry_fixup
    for lb in range(1, nblk):
        if not flagged[lb]:
            continue
        total = IDENT
        for m in maps[:lb]:
            total = compose(m, total)
        c = apply(total, 0)
        j = lb * WB
        while c and j < min(n, (lb + 1) * WB):
            v = out[j] + c
            out[j], c = v % M, v >> 32
            j += 1
    return out, flagged


def _sums(rng, n, kind):
    if kind == "random":
        return [rng.randrange(6 * M) for _ in range(n)]
    if kind == "ones":  # every sum 2^32 - 1: every word passes a carry on, no window resolves
        return [M - 1] * n
    if kind == "near":  # low halves within 7 of 2^32
        return [rng.randrange(5) * M + M - 1 - rng.randrange(7) for _ in range(n)]
    # mixture: long all-ones stretches inside random ones
    out = []
    while len(out) < n:
        run = rng.randrange(1, 5000)
        out += _sums(rng, run, rng.choice(["random", "ones"]))
    return out[:n]


@pytest.mark.parametrize("kind", ["random", "ones", "near", "mixed"])
@pytest.mark.parametrize("n", [1, 7, 2016, 2017, 2016 * 3 + 5, 20000])
def test_blocked_carry_equals_sequential(kind, n):
    rng = random.Random(hash((kind, n)) & 0xFFFF)
    sums = _sums(rng, n, kind)
    for j in range(max(0, n - 4), n):  # a product's top words absorb every carry
        sums[j] = 0
    want = sequential(sums)
    got, flagged = blocked(sums)
    assert got == want
    if kind == "random" and n > WB:
        assert not any(flagged)  # random sums always resolve
    if kind == "ones":  # every block whose window lies in the all-ones words takes the fixup path
        assert all(flagged[lb] for lb in range(1, len(flagged)) if lb * WB <= n - 4)


def test_planted_mistake_fails():
    rng = random.Random(3)
    # A carry of 1 entering a long run of 2^32 - 1 sums: every block's true
    # carry-in is 1, while every window's map is c -> [c >= 1].
    sums = [2 * M] + [M - 1] * 8995 + [0, 0, 0, 0]
    want = sequential(sums)
    assert blocked(sums)[0] == want
    assert blocked(sums, mistake="window_always_resolved")[0] != want
